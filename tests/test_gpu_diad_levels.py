"""Wall-refined DIAD levels on the device: transfers, level recursion, drag.

A sphere channel is built on the device (psl_diad_build; its tables equal the
oracle's bit for bit, tests/test_gpu_diad_builder.py), each level becomes an
esotwist grid (psl_diad_create_level) with the DIAD sphere driver's flags, and
the run is compared with the oracle's recursion over the same tables
(tests/eso_levels.py): populations of every level after each block of coarse
steps within 1e-5 relative (FP32 storage) and the FP64 momentum exchange on
the finest level.  The coarse<->fine communication alone is compared on a
random state, and FP16 storage is held to the FP32 oracle with a bound on the
absolute error.
"""
import numpy as np
import pytest

import oracle as O
from tests import eso_levels as E

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b), 1e-3)))


def _oracle_levels(P, lv, dtype=np.float32, cs=0.1):
    out = []
    for l in range(lv.levels):
        d = lv.level(l)
        info = d["info"]
        dd = dict(d, nx=info.cellCountX, ny=info.cellCountY, nz=info.cellCountZ)
        flags = cases.diad_sphere_flags(d["ijk"], d["bounceback"], (dd["nx"], dd["ny"], dd["nz"]), l)
        out.append(E.OracleLevel(dd, flags, cases.channel_given_values(0.04), info.nu, cs, dtype=dtype))
    return out


def _setup(P, storage, n0=16, levels=3, ux0=0.0):
    lv, grids = cases.diad_sphere(P, storage, n0=n0, levels=levels, span=1, radius=2.0)
    ol = _oracle_levels(P, lv)
    for g, o in zip(grids, ol):
        if ux0:
            g.fill_equilibrium(ux=ux0)
        o.set(g.get_f())
    counts = [g.n for g in grids]
    print(f"[diad-levels] storage={storage} cells per level {counts}", flush=True)
    return lv, grids, ol


def test_diad_levels_match_oracle(psl):
    P = psl
    # started from uniform flow so that the sphere carries a load from the first step
    lv, grids, ol = _setup(P, "f32", ux0=0.04)
    for block in range(3):
        P.updateGrid(grids, 2)
        for _ in range(2):
            E.update_grid(ol)
        # the recursion's transfers read the interface records the updates wrote
        recs = [g.interface_records()[0] for g in grids]
        print(f"  interface records per level {recs}", flush=True)
        assert all(r > 0 for r in recs)
        for l, (g, o) in enumerate(zip(grids, ol)):
            assert g.esotwistFlipper == o.flip
            err = _rel_err(g.get_f(), o.get())
            print(f"  after {2 * (block + 1)} coarse steps, level {l}: max rel err {err:.3e}", flush=True)
            assert err < RTOL
    fx = grids[-1].momentum_exchange()
    fo = O.eso_momentum_exchange(ol[-1].f, ol[-1].nbr7, ol[-1].flip, ol[-1].flags)
    print(f"  momentum exchange GPU {fx} oracle {fo}", flush=True)
    assert abs(fx[0] - fo[0]) <= 1e-4 * abs(fo[0])
    assert np.all(np.abs(fx - fo) <= 1e-4 * abs(fo[0]))
    # FP64 from the GPU's own populations: bounceback reverses the momentum, so
    # rho (u - u_bounced) = 2 sum_q c_q f_q per bounceback cell
    f = grids[-1].get_f().astype(np.float64)
    bb = ol[-1].flags == cases.BOUNCEBACK
    cx, cy, cz = (np.asarray(c, np.float64) for c in O.velocity_set())
    m = np.stack([2.0 * (c[:, None] * f[:, bb]).sum(axis=0) for c in (cx, cy, cz)])
    ref = m.sum(axis=1)
    # FP32 rounding bound of the per-cell sums: 27 terms at 2^-24 relative each
    scale = np.array([2.0 * (np.abs(c)[:, None] * np.abs(f[:, bb])).sum() for c in (cx, cy, cz)])
    print(f"  FP64 from the populations {ref}", flush=True)
    assert np.all(np.abs(fx - ref) <= 4e-6 * scale)


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_interface_records_hold_each_update(psl, storage):
    """Every record equals, word for word, what its cell's slots hold right
    after a local update of its level; a state set from the host is never read
    from records."""
    P = psl
    lv, grids = cases.diad_sphere(P, storage, n0=16, levels=3, span=1, radius=2.0)
    for g in grids:
        g.fill_equilibrium(ux=0.04)
    P.updateGrid(grids, 1)
    for l, g in enumerate(grids):
        n, _ = g.interface_records()
        P.applyStreaming(g)
        P.applyLocalCellUpdate(g)
        bad = g.verify_interface_records()
        print(f"[records] {storage} level {l}: {n} records, {bad} differ after an update", flush=True)
        assert n > 0
        assert g.interface_records() == (n, True)
        assert bad == 0
        g.set_f(g.get_f())
        assert g.interface_records() == (n, False)


def test_diad_communication_random_state(psl):
    P = psl
    lv, grids, ol = _setup(P, "f32", n0=12, levels=2)
    rng = np.random.default_rng(3)
    for g, o in zip(grids, ol):
        ru = np.column_stack([1 + 0.01 * rng.standard_normal(g.n), 0.03 * rng.standard_normal((g.n, 3))])
        f = (O.feq(ru.astype(np.float32)) + np.float32(1e-3) * rng.standard_normal((g.n, 27)).astype(np.float32)).T
        f = np.ascontiguousarray(f, np.float32)
        g.set_f(f)
        o.set(f)
    P.applyCoarseFineGridCommunication(grids[0], grids[1])
    E.communicate(ol[0], ol[1])
    for l in range(2):
        err = _rel_err(grids[l].get_f(), ol[l].get())
        print(f"  communication level {l}: max rel err {err:.3e}", flush=True)
        assert err < RTOL


def test_diad_levels_f16_storage(psl):
    P = psl
    lv, grids, ol = _setup(P, "f16")
    P.updateGrid(grids, 4)
    for _ in range(4):
        E.update_grid(ol)
    for l, (g, o) in enumerate(zip(grids, ol)):
        err = float(np.max(np.abs(g.get_f().astype(np.float64) - o.get())))
        print(f"  f16 level {l}: max abs err {err:.3e}", flush=True)
        assert err < 2e-3


def test_diad_level_grid_rejects_mixed_levels(psl):
    P = psl
    lv, grids = cases.diad_sphere(P, "f32", n0=12, levels=2, span=1, radius=2.0)
    other = P.DIADGridStruct.from_levels(lv, 1, "f16")
    with pytest.raises(P.PslError):
        P.applyCoarseFineGridCommunication(grids[0], other)


@pytest.mark.parametrize("flip_first", [False, True])
def test_nonreflective_outlet_matches_oracle(psl, flip_first):
    """applyNonReflectiveOutletZ on a level built with an outlet rule (outlet
    list and -k neighbours come from the builder), on a random state, then one
    more streaming + local update."""
    P = psl
    from tests.test_gpu_diad_builder import _gpu_build
    from tests.test_diad_levels import outlet_level
    base = dict(res=1.0, ox=-5.5, oy=-5.5, oz=-11.5, nu=0.02, dt=1e-3, nx=16, ny=12, nz=24)
    rules = P.rules(cases.diad_sphere_rules(P, 1.6, 2),
                    P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(0, -1, 0, -1, -1, -1), levels=[0]))
    lv = _gpu_build(P, base, 2, 1, [], rules, nu_phys=2e-5)
    d = outlet_level()
    assert np.array_equal(lv.array(0, "outlet"), d["outlet"])
    g = P.DIADGridStruct.from_levels(lv, 0, "f32")
    if flip_first:
        P.applyStreaming(g)
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    rng = np.random.default_rng(7)
    ru = np.column_stack([1 + 0.01 * rng.standard_normal(g.n), 0.03 * rng.standard_normal((g.n, 3))])
    f = np.ascontiguousarray((O.feq(ru.astype(np.float32)) + np.float32(1e-3) *
                              rng.standard_normal((g.n, 27)).astype(np.float32)).T, np.float32)
    g.set_f(f)
    raw = np.zeros((27, g.n), np.float32)
    E.eso_set_all(raw, nbr7, flip_first, f)
    P.applyNonReflectiveOutletZ(g)
    O.eso_outlet_z(raw, nbr7, flip_first, d["outlet"], d["nbr"][9])
    err = _rel_err(g.get_f(), E.eso_get_all(raw, nbr7, flip_first))
    print(f"  outlet (flipper {flip_first}): max rel err {err:.3e} over {d['outlet'].size} outlet cells", flush=True)
    assert err < RTOL
    # the next step reads what the operator wrote
    flags = cases.diad_sphere_flags(d["ijk"], d["bounceback"], (d["nx"], d["ny"], d["nz"]), 0)
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(0.04))
    P.applyStreaming(g)
    P.applyLocalCellUpdate(g)
    info = O.make_info(d["nx"], d["ny"], d["nz"], nu=0.02, cs=0.0)
    O.esotwist_update(raw, nbr7, np.ascontiguousarray(d["ijk"], np.int32), info, flags,
                      cases.channel_given_values(0.04).reshape(-1), not flip_first)
    err = _rel_err(g.get_f(), E.eso_get_all(raw, nbr7, not flip_first))
    print(f"  after one step: max rel err {err:.3e}", flush=True)
    assert err < RTOL


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_moving_bounceback_matches_oracle(psl, storage):
    """Moving-wall cells (flagged bounceback, velocities per cell) over three
    steps of streaming + local update against the oracle's update followed by
    applyMovingBounceback on the same cells."""
    P = psl
    from tests.test_gpu_diad_builder import _gpu_build
    from tests.test_diad_levels import outlet_level
    base = dict(res=1.0, ox=-5.5, oy=-5.5, oz=-11.5, nu=0.02, dt=1e-3, nx=16, ny=12, nz=24)
    rules = P.rules(cases.diad_sphere_rules(P, 1.6, 2),
                    P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(0, -1, 0, -1, -1, -1), levels=[0]))
    lv = _gpu_build(P, base, 2, 1, [], rules, nu_phys=2e-5)
    d = outlet_level()
    g = P.DIADGridStruct.from_levels(lv, 0, storage)
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    rng = np.random.default_rng(11)
    flags = cases.diad_sphere_flags(d["ijk"], d["bounceback"], (d["nx"], d["ny"], d["nz"]), 0)
    interior = np.flatnonzero(flags == cases.FLUID)
    picked = rng.choice(interior, 250, replace=False)
    cells = np.sort(picked[:150]).astype(np.int32)
    flags[picked[150:]] = cases.BOUNCEBACK  # static walls beside the moving ones
    walls_only = flags.copy()  # the reference's bounceback marker leaves moving walls out
    flags[cells] = cases.BOUNCEBACK
    u = (0.02 * rng.standard_normal((cells.size, 3))).astype(np.float32)
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(0.04))
    g.set_moving_bounceback(cells, u)
    g.fill_equilibrium(ux=0.03)
    raw = np.zeros((27, g.n), np.float32)
    E.eso_set_all(raw, nbr7, False, g.get_f())
    info = O.make_info(d["nx"], d["ny"], d["nz"], nu=0.02, cs=0.0)
    ijk = np.ascontiguousarray(d["ijk"], np.int32)
    flip = False
    for step in range(3):
        P.applyStreaming(g)
        P.applyLocalCellUpdate(g)
        flip = not flip
        O.esotwist_update(raw, nbr7, ijk, info, flags, cases.channel_given_values(0.04).reshape(-1), flip)
        O.eso_moving_bounceback(raw, nbr7, flip, cells, u)
        ref = E.eso_get_all(raw, nbr7, flip)
        got = g.get_f()
        if storage == "f32":
            err = _rel_err(got, ref)
            print(f"  step {step + 1}: max rel err {err:.3e}", flush=True)
            assert err < RTOL
        else:
            err = float(np.max(np.abs(got.astype(np.float64) - ref)))
            print(f"  step {step + 1}: max abs err {err:.3e} (f16 storage vs FP32 oracle)", flush=True)
            assert err < 2e-3
    # the drag leaves the moving-wall cells out
    fx = g.momentum_exchange()
    fo = O.eso_momentum_exchange(raw, nbr7, flip, walls_only)
    print(f"  momentum exchange without moving walls: GPU {fx} oracle {fo}", flush=True)
    assert np.abs(fo).max() > 0
    tol = (1e-4 if storage == "f32" else 2e-2) * np.abs(fo).max()
    assert np.all(np.abs(fx - fo) <= tol)


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_forced_velocity_matches_oracle(psl, storage):
    """Forced-velocity cells (flagged bounceback, target velocities per cell)
    beside moving walls and static walls, over three steps (streaming + local
    update, then the level recursion on one level) against the oracle's update
    followed by its moving-wall and forced-cell operators; then the forcing
    sums (force and torque) and the drag, which leaves both lists out."""
    P = psl
    from tests.test_gpu_diad_builder import _gpu_build
    from tests.test_diad_levels import outlet_level
    base = dict(res=1.0, ox=-5.5, oy=-5.5, oz=-11.5, nu=0.02, dt=1e-3, nx=16, ny=12, nz=24)
    rules = P.rules(cases.diad_sphere_rules(P, 1.6, 2),
                    P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(0, -1, 0, -1, -1, -1), levels=[0]))
    lv = _gpu_build(P, base, 2, 1, [], rules, nu_phys=2e-5)
    d = outlet_level()
    g = P.DIADGridStruct.from_levels(lv, 0, storage)
    g.set_smagorinsky(0.1)
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    rng = np.random.default_rng(21)
    flags = cases.diad_sphere_flags(d["ijk"], d["bounceback"], (d["nx"], d["ny"], d["nz"]), 0)
    interior = np.flatnonzero(flags == cases.FLUID)
    picked = rng.choice(interior, 400, replace=False)
    forced = np.sort(picked[:200]).astype(np.int32)
    moving = np.sort(picked[200:300]).astype(np.int32)
    flags[picked[300:]] = cases.BOUNCEBACK  # static walls
    walls_only = flags.copy()
    flags[forced] = cases.BOUNCEBACK
    flags[moving] = cases.BOUNCEBACK
    ut = (0.03 * rng.standard_normal((forced.size, 3))).astype(np.float32)
    um = (0.02 * rng.standard_normal((moving.size, 3))).astype(np.float32)
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(0.04))
    g.set_moving_bounceback(moving, um)
    g.set_forced_velocity(forced, ut)
    with pytest.raises(P.PslError):  # a cell may not be in both lists
        g.set_forced_velocity(np.concatenate([forced, moving[:1]]), np.vstack([ut, um[:1]]))
    # listed cells must stay flagged bounceback: flags that make them fluid are
    # refused (and the grid keeps its flags), as is a fluid cell in a list
    with pytest.raises(P.PslError):
        g.set_flags(walls_only)
    with pytest.raises(P.PslError):
        g.set_forced_velocity(np.concatenate([forced, np.setdiff1d(interior, picked)[:1]]), np.vstack([ut, ut[:1]]))
    with pytest.raises(ValueError):  # three velocity components per cell
        g.set_forced_velocity(forced, ut[:-1])
    g.set_forced_velocity(forced, ut)
    g.fill_equilibrium(ux=0.03)
    raw = np.zeros((27, g.n), np.float32)
    E.eso_set_all(raw, nbr7, False, g.get_f())
    info = O.make_info(d["nx"], d["ny"], d["nz"], nu=0.02, cs=0.1)
    ijk = np.ascontiguousarray(d["ijk"], np.int32)
    flip = False
    for step in range(3):
        if step == 1:
            P.updateGrid([g], 1)
        else:
            P.applyStreaming(g)
            P.applyLocalCellUpdate(g)
        flip = not flip
        O.esotwist_update(raw, nbr7, ijk, info, flags, cases.channel_given_values(0.04).reshape(-1), flip)
        O.eso_moving_bounceback(raw, nbr7, flip, moving, um)
        O.eso_forced_velocity(raw, nbr7, flip, forced, ut, np.float32(0.02), np.float32(0.1))
        ref = E.eso_get_all(raw, nbr7, flip)
        got = g.get_f()
        if storage == "f32":
            err = _rel_err(got, ref)
            print(f"  step {step + 1}: max rel err {err:.3e}", flush=True)
            assert err < RTOL
            # the forced cells leave the update at their targets
            u = O.rho_u(np.ascontiguousarray(got[:, forced].T.astype(np.float64)))[:, 1:]
            assert np.max(np.abs(u - ut)) < 5e-6
        else:
            err = float(np.max(np.abs(got.astype(np.float64) - ref)))
            print(f"  step {step + 1}: max abs err {err:.3e} (f16 storage vs FP32 oracle)", flush=True)
            assert err < 2e-3
    lvl, _ = lv.info(0)
    oinfo = dict(res=lvl.res, ox=lvl.ox, oy=lvl.oy, oz=lvl.oz, dt=lvl.dtPhys)
    fo = O.eso_forcing_sum(raw, nbr7, flip, forced, ut, ijk, oinfo, 1000.0)
    fg = g.forcing(1000.0)
    print(f"  forcing GPU {fg} oracle {fo}", flush=True)
    assert np.abs(fo).max() > 0
    tol = (1e-4 if storage == "f32" else 2e-2) * np.abs(fo[:3]).max()
    assert np.all(np.abs(fg[:3] - fo[:3]) <= tol)
    ttol = (1e-4 if storage == "f32" else 2e-2) * np.abs(fo[3:]).max()
    assert np.all(np.abs(fg[3:] - fo[3:]) <= ttol)
    assert P.getTorque(g, 1000.0) == fg[5]
    fx = g.momentum_exchange()
    mo = O.eso_momentum_exchange(raw, nbr7, flip, walls_only)
    print(f"  drag without moving and forced cells: GPU {fx} oracle {mo}", flush=True)
    tol = (1e-4 if storage == "f32" else 2e-2) * np.abs(mo).max()
    assert np.all(np.abs(fx - mo) <= tol)


def _box_triangles(lo, hi):
    """12 outward-wound triangles of an axis-aligned box."""
    x0, y0, z0 = lo
    x1, y1, z1 = hi
    v = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    faces = [(0, 2, 1), (0, 3, 2), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4),
             (2, 3, 7), (2, 7, 6), (1, 2, 6), (1, 6, 5), (3, 0, 4), (3, 4, 7)]
    return np.array([np.concatenate([v[a], v[b], v[c]]) for a, b, c in faces], np.float32)


def test_rotating_stl_markers_match_oracle(psl):
    """The M35 impeller loop: an STL blade rotated about z (rotateSTLAlongZ),
    then ApplyMarkersInsideSTL on each DIAD level's cells, bit for bit against
    the oracle's dense-box voxelizer sampled at the level's (i, j, k).  The
    rotation matches the closed form, and turning back restores the blade."""
    P = psl
    lv, grids = cases.diad_sphere(P, "f32", n0=16, levels=2, span=1, radius=2.0)
    tri0 = _box_triangles((-4.1, -0.7, -2.2), (4.1, 0.7, 2.2))
    stl = P.STLStruct.from_triangles(tri0)
    total = 0
    for step, angle in enumerate((0.0, 0.3, 0.3, 0.45)):
        if angle:
            P.rotateSTLAlongZ(stl, angle)
        tri = stl.triangles()
        th = sum((0.0, 0.3, 0.3, 0.45)[: step + 1])
        c, s = np.cos(th), np.sin(th)
        v0 = tri0.reshape(-1, 3).astype(np.float64)
        want_v = np.column_stack([v0[:, 0] * c - v0[:, 1] * s, v0[:, 0] * s + v0[:, 1] * c, v0[:, 2]])
        assert np.max(np.abs(tri.reshape(-1, 3) - want_v)) < 1e-5
        b = stl.bounds
        assert np.allclose(b[:3], want_v.min(0), atol=1e-5) and np.allclose(b[3:], want_v.max(0), atol=1e-5)
        for l, g in enumerate(grids):
            info, _ = lv.info(l)
            got = g.markers_inside_stl(stl)
            dense = O.markers_inside_stl(tri, (0, 0, 0), (info.ox, info.oy, info.oz), info.res,
                                         info.cellCountX, info.cellCountY, info.cellCountZ)
            i, j, k = (a.astype(np.int64) for a in lv.array(l, "ijk"))
            want = dense[(k * info.cellCountY + j) * info.cellCountX + i]
            print(f"  angle {th:.2f} level {l}: {int(got.sum())} cells inside", flush=True)
            assert np.array_equal(got, want)
            total += int(got.sum())
    assert total > 0
    P.rotateSTLAlongZ(stl, -1.05)
    assert np.max(np.abs(stl.triangles() - tri0)) < 1e-5


def test_forced_and_moving_cells_through_the_level_recursion(psl):
    """Forced cells on the finest level and moving walls on the middle one,
    advanced by the three-level recursion (updates, both list kernels and the
    coarse<->fine transfers) against the oracle's recursion with the same lists:
    populations of every level within 1e-5 relative after each coarse step."""
    P = psl
    lv, grids, ol = _setup(P, "f32", ux0=0.04)
    rng = np.random.default_rng(17)
    for l, lists in ((2, "forced"), (1, "moving")):
        o = ol[l]
        cand = np.flatnonzero(o.flags == cases.FLUID)
        cells = np.sort(rng.choice(cand, min(120, cand.size // 10), replace=False)).astype(np.int32)
        o.flags[cells] = cases.BOUNCEBACK
        grids[l].set_flags(o.flags)
        vel = (0.03 * rng.standard_normal((cells.size, 3))).astype(np.float32)
        if lists == "forced":
            grids[l].set_forced_velocity(cells, vel)
            o.forced = (cells, vel)
        else:
            grids[l].set_moving_bounceback(cells, vel)
            o.moving = (cells, vel)
        print(f"  level {l}: {cells.size} {lists} cells", flush=True)
    for block in range(2):
        P.updateGrid(grids, 1)
        E.update_grid(ol)
        for l, (g, o) in enumerate(zip(grids, ol)):
            err = _rel_err(g.get_f(), o.get())
            print(f"  after {block + 1} coarse steps, level {l}: max rel err {err:.3e}", flush=True)
            assert err < RTOL
