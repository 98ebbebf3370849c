"""DIAD level transfers in the oracle (include/gridRefinementFunctions.h:540-708): CPU pins.

On a two-level wall-refined sphere channel built by the oracle's builder:
a uniform rest equilibrium passes both transfers unchanged; every coarse
fine-to-coarse cell receives rescaleF(mean of its eight children, fine->coarse)
and every child of a coarse-to-fine cell receives rescaleF(coarse, coarse->fine),
where the children are looked up by their (2i+a, 2j+b, 2k+c) coordinates in the
fine cell table, not through the esotwist neighbour maps the transfers use.
FP64 instantiation, so the pins are tight.
"""
import numpy as np
import pytest

import oracle as O
from tests import eso_levels as E

import periodicshiftlbm_amd as P
from periodicshiftlbm_amd import cases


@pytest.fixture(scope="module")
def two_levels():
    base = dict(res=1.0, ox=-5.5, oy=-5.5, oz=-5.5, nu=0.02, dt=1e-3, nx=24, ny=12, nz=12)
    lv = O.diad_build(base, 2, 1, 2e-5, (), cases.diad_sphere_rules(P, 1.6, 2))
    bc = cases.channel_given_values(0.04)
    levels = []
    for l, d in enumerate(lv):
        flags = cases.diad_sphere_flags(d["ijk"], d["bounceback"], (d["nx"], d["ny"], d["nz"]), l)
        levels.append(E.OracleLevel(d, flags, bc, d["nu"], 0.1, dtype=np.float64))
    assert lv[0]["f2c_write"].size > 0 and lv[0]["c2f_read"].size > 0
    return lv, levels


def _children(d_c, d_f, cells):
    i, j, k = (a.astype(np.int64) for a in d_c["ijk"])
    fi, fj, fk = (a.astype(np.int64) for a in d_f["ijk"])
    key = (fk * d_f["ny"] + fj) * d_f["nx"] + fi
    order = np.argsort(key)
    out = []
    for a in (0, 1):
        for b in (0, 1):
            for c in (0, 1):
                want = ((2 * k[cells] + c) * d_f["ny"] + 2 * j[cells] + b) * d_f["nx"] + 2 * i[cells] + a
                pos = np.searchsorted(key[order], want)
                assert np.all(key[order][pos] == want)
                out.append(order[pos])
    return np.array(out)  # 8 x m


def test_uniform_equilibrium_passes_unchanged(two_levels):
    lv, (c, f) = two_levels
    # a rest state (the polynomial equilibrium at u != 0 is not a fixed point of
    # rescaleF: its higher central moments differ from the product form)
    e = O.feq(np.array([1.02, 0.0, 0.0, 0.0]))
    for L, flip in ((c, False), (f, True)):
        L.flip = flip
        L.set(np.repeat(e[:, None], L.n, axis=1))
    E.communicate(c, f)
    for L in (c, f):
        assert np.max(np.abs(L.get() - e[:, None])) < 1e-14


def test_transfers_route_to_children(two_levels):
    lv, (c, f) = two_levels
    rng = np.random.default_rng(5)
    for L in (c, f):
        L.flip = bool(rng.integers(2))
        ru = np.column_stack([1 + 0.01 * rng.standard_normal(L.n), 0.03 * rng.standard_normal((L.n, 3))])
        L.set((O.feq(ru) + 1e-3 * rng.standard_normal((L.n, 27))).T)
    fc0, ff0 = c.get(), f.get()
    E.communicate(c, f)
    fc1, ff1 = c.get(), f.get()
    # fine -> coarse
    w, r = lv[0]["f2c_write"], lv[0]["f2c_read"]
    ch = _children(lv[0], lv[1], w)
    assert np.array_equal(ch[0], r)
    mean = ff0[:, ch].mean(axis=1).T  # m x 27
    want = O.rescale(np.ascontiguousarray(mean), False)
    assert np.max(np.abs(fc1[:, w].T - want)) < 1e-14
    # coarse -> fine
    w2, r2 = lv[0]["c2f_write"], lv[0]["c2f_read"]
    ch2 = _children(lv[0], lv[1], r2)
    assert np.array_equal(ch2[0], w2)
    want2 = O.rescale(np.ascontiguousarray(fc0[:, r2].T), True)
    for a in range(8):
        assert np.max(np.abs(ff1[:, ch2[a]].T - want2)) < 1e-14
    # everything else untouched
    others = np.setdiff1d(np.arange(c.n), w)
    assert np.array_equal(fc1[:, others], fc0[:, others])
    touched = np.unique(ch2)
    others_f = np.setdiff1d(np.arange(f.n), touched)
    assert np.array_equal(ff1[:, others_f], ff0[:, others_f])


# ------------------------------------------------ non-reflective z outlet
QZ = [3, 7, 10, 13, 18, 19, 22, 23, 25]
INV_SQRT3 = np.float64(np.float32(0.577350269))
ONE_MINUS = np.float64(np.float32(1.0) - np.float32(0.577350269))


def outlet_level():
    """Level 0 of a two-level sphere box with an outlet rule on its last k plane;
    the refined region stays clear of the two last planes, so every outlet
    cell's -k neighbour is the cell below it (the operator's reads and writes
    then touch disjoint slots, as in the drivers' geometries)."""
    base = dict(res=1.0, ox=-5.5, oy=-5.5, oz=-11.5, nu=0.02, dt=1e-3, nx=16, ny=12, nz=24)
    rules = P.rules(cases.diad_sphere_rules(P, 1.6, 2),
                    P.marker_rule(P.RULE_OUTLET, P.SHAPE_INDEX_BOX, idx=(0, -1, 0, -1, -1, -1), levels=[0]))
    return O.diad_build(base, 2, 1, 2e-5, (), rules)[0]


def _read_slots(d, flip):
    """(fi, slot) of the read maps of parity `flip` for every (q, cell)."""
    R, _ = O.esotwist_tables()
    s = np.vstack([np.arange(d["n"])[None], d["nbr"][:7].astype(np.int64)])
    return [((q if flip else E.OPP[q]), s[R[q]]) for q in range(27)]


@pytest.mark.parametrize("flip", [False, True])
def test_outlet_extrapolates_uniform_and_linear_states(flip):
    d = outlet_level()
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    kmin = d["nbr"][9]
    out = d["outlet"]
    i, j, k = d["ijk"]
    assert out.size == d["nx"] * d["ny"] and np.all(k[out] == d["nz"] - 1) and np.all(k[kmin[out]] == d["nz"] - 2)
    rng = np.random.default_rng(1)
    # uniform state: the extrapolation returns the state; nothing else changes
    e = rng.random(27)
    raw = np.zeros((27, d["n"]))
    E.eso_set_all(raw, nbr7, flip, np.repeat(e[:, None], d["n"], axis=1))
    raw0 = raw.copy()
    O.eso_outlet_z(raw, nbr7, flip, out, kmin)
    rs = _read_slots(d, not flip)
    touched = np.zeros_like(raw, bool)
    for q in QZ:
        fi, sl = rs[q]
        assert np.max(np.abs(raw[fi, sl[out]] - e[q])) < 1e-15
        touched[fi, sl[out]] = True
    assert np.array_equal(raw[~touched], raw0[~touched])
    # linear in k: invSqrt3 f(-k neighbour) + (1 - invSqrt3) f(cell)
    a, b = rng.random(27), rng.random(27)
    f = a[:, None] + b[:, None] * k[None, :].astype(np.float64)
    raw = np.zeros((27, d["n"]))
    E.eso_set_all(raw, nbr7, flip, f)
    O.eso_outlet_z(raw, nbr7, flip, out, kmin)
    for q in QZ:
        fi, sl = rs[q]
        want = INV_SQRT3 * (a[q] + b[q] * k[kmin[out]]) + ONE_MINUS * (a[q] + b[q] * k[out])
        assert np.max(np.abs(raw[fi, sl[out]] - want)) < 1e-12


# ------------------------------------------------------------ moving walls
@pytest.mark.parametrize("flip", [False, True])
def test_moving_bounceback_pins(flip):
    """A wall at rest leaves every stored value as it was (the implicit
    bounceback of a skipped cell); a moving wall conserves the cell's mass and
    turns its momentum m into -m + 2 u (sum_q 6 w_q c_q c_q = 2 I)."""
    d = outlet_level()
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    rng = np.random.default_rng(4 + flip)
    raw = rng.random((27, d["n"]))
    cells = np.sort(rng.choice(d["n"], 200, replace=False)).astype(np.int32)
    before = raw.copy()
    O.eso_moving_bounceback(raw, nbr7, flip, cells, np.zeros((cells.size, 3), np.float32))
    assert np.array_equal(raw, before)
    u = (0.05 * rng.standard_normal((cells.size, 3))).astype(np.float32)
    R, _ = O.esotwist_tables()
    s = np.vstack([np.arange(d["n"])[None], d["nbr"][:7].astype(np.int64)])
    f_in = np.stack([raw[q if flip else E.OPP[q], s[R[q]][cells]] for q in range(27)])  # read maps
    O.eso_moving_bounceback(raw, nbr7, flip, cells, u)
    f_out = E.eso_get_all(raw, nbr7, flip)[:, cells]                                    # write maps
    c = np.array(O.velocity_set(), np.float64)                                          # 3 x 27
    assert np.max(np.abs(f_out.sum(0) - f_in.sum(0))) < 1e-12
    m_in, m_out = c @ f_in, c @ f_out
    assert np.max(np.abs(m_out - (-m_in + 2.0 * u.T))) < 1e-7


# ------------------------------------------------------- forced velocity
def test_forcing_collision_pins():
    """The forcing collision conserves mass and adds the force g to the
    momentum (first-order central moments -g/2 about u' = u + g/(2 rho),
    restored as +g/2); with g = 0 it is the unforced collision bit for bit.
    A dropped half-force shift, a sign flip of the restored first-order moments
    or a velocity left unshifted each break the momentum balance."""
    rng = np.random.default_rng(8)
    c = np.array(O.velocity_set(), np.float64)
    for cs in (0.0, 0.1):
        for _ in range(20):
            ru = np.array([1 + 0.02 * rng.standard_normal(), *(0.05 * rng.standard_normal(3))])
            f = O.feq(ru) + 1e-3 * rng.standard_normal(27)
            g = 2e-3 * rng.standard_normal(3)
            out = O.collide_forced(f, 0.02, cs, g)
            assert abs(out.sum() - f.sum()) < 1e-14
            assert np.max(np.abs(c @ out - (c @ f + g))) < 1e-14
            assert np.array_equal(O.collide_forced(f, 0.02, cs, np.zeros(3)), O.collide(f.copy(), 0.02, cs))


@pytest.mark.parametrize("flip", [False, True])
def test_forced_cells_reach_their_target(flip):
    """With g = rho (u_target - u) the momentum after the forcing collision is
    rho u_target: every forced cell leaves the update at its target velocity
    with its mass; slots of other cells are untouched."""
    d = outlet_level()
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    rng = np.random.default_rng(12 + flip)
    ru = np.column_stack([1 + 0.01 * rng.standard_normal(d["n"]), 0.03 * rng.standard_normal((d["n"], 3))])
    f = (O.feq(ru) + 1e-3 * rng.standard_normal((d["n"], 27))).T
    raw = np.zeros((27, d["n"]))
    E.eso_set_all(raw, nbr7, not flip, f)  # the previous update's writes = this step's reads
    cells = np.sort(rng.choice(d["n"], 150, replace=False)).astype(np.int32)
    ut = (0.04 * rng.standard_normal((cells.size, 3))).astype(np.float32)
    before = raw.copy()
    R, _ = O.esotwist_tables()
    s = np.vstack([np.arange(d["n"])[None], d["nbr"][:7].astype(np.int64)])
    f_in = np.stack([raw[q if flip else E.OPP[q], s[R[q]][cells]] for q in range(27)])  # read maps
    O.eso_forced_velocity(raw, nbr7, flip, cells, ut, 0.02, 0.1)
    out = E.eso_get_all(raw, nbr7, flip)[:, cells]
    m = O.rho_u(np.ascontiguousarray(out.T))
    assert np.max(np.abs(m[:, 0] - f_in.sum(0))) < 1e-13
    assert np.max(np.abs(m[:, 1:] - ut)) < 1e-13
    touched = np.zeros_like(raw, bool)
    for q in range(27):
        touched[q if flip else E.OPP[q], s[R[q]][cells]] = True
    assert np.array_equal(raw[~touched], before[~touched])


def test_forcing_sum_rigid_rotation():
    """Cells at rest (rho = 1) driven towards a rigid rotation u_t = w e_z x r:
    g = u_t, so the force is scale * sum u_t and the torque about z is
    scale * w * sum (x^2 + y^2) / 1000, with scale = rho_phys (res/1000)^4 / dt^2."""
    d = outlet_level()
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    raw = np.zeros((27, d["n"]))
    e = O.feq(np.array([1.0, 0.0, 0.0, 0.0]))
    E.eso_set_all(raw, nbr7, False, np.repeat(e[:, None], d["n"], axis=1))
    rng = np.random.default_rng(2)
    cells = np.sort(rng.choice(d["n"], 300, replace=False)).astype(np.int32)
    info = dict(res=100.0, ox=-550.0, oy=-550.0, oz=-1150.0, dt=0.01)
    i, j, k = (a[cells].astype(np.float64) for a in d["ijk"])
    x, y, z = (a * info["res"] + o for a, o in ((i, info["ox"]), (j, info["oy"]), (k, info["oz"])))
    w = 1e-5
    ut = np.column_stack([-w * y, w * x, np.zeros_like(x)]).astype(np.float32)
    got = O.eso_forcing_sum(raw, nbr7, False, cells, ut, d["ijk"], info, 1000.0)
    scale = 1000.0 * (info["res"] / 1000.0) ** 4 / info["dt"] ** 2
    g = ut.astype(np.float64) * scale
    assert np.allclose(got[:3], g.sum(0), rtol=1e-6, atol=1e-9)
    assert np.isclose(got[5], scale * w * np.sum(x * x + y * y) / 1000.0, rtol=1e-6)
    assert np.isclose(got[3], np.sum(-z * g[:, 1]) / 1000.0, rtol=1e-6, atol=1e-9)
    assert np.isclose(got[4], np.sum(z * g[:, 0]) / 1000.0, rtol=1e-6, atol=1e-9)


@pytest.mark.parametrize("flip", [False, True])
def test_forcing_sum_reads_the_next_step(flip):
    """The forcing sum sees the populations the next update reads (read maps of
    the other parity), not the cell's own last writes: on a random state it
    equals scale * sum rho (u_t - u) over that gather, and differs from the
    same sum over the write maps."""
    d = outlet_level()
    nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
    rng = np.random.default_rng(30 + flip)
    raw = 0.03 + 0.01 * rng.random((27, d["n"]))
    cells = np.sort(rng.choice(d["n"], 100, replace=False)).astype(np.int32)
    ut = (0.02 * rng.standard_normal((cells.size, 3))).astype(np.float32)
    info = dict(res=1.0, ox=0.0, oy=0.0, oz=0.0, dt=1.0)
    got = O.eso_forcing_sum(raw, nbr7, flip, cells, ut, d["ijk"], info, 1.0)
    scale = 1.0 * (1.0 / 1000.0) ** 4
    R, _ = O.esotwist_tables()
    s = np.vstack([np.arange(d["n"])[None], d["nbr"][:7].astype(np.int64)])
    nxt = np.stack([raw[E.OPP[q] if flip else q, s[R[q]][cells]] for q in range(27)])
    own = E.eso_get_all(raw, nbr7, flip)[:, cells]
    c = np.array(O.velocity_set(), np.float64)

    def force(f):
        rho = f.sum(0)
        return (scale * (rho[:, None] * ut - (c @ f).T)).sum(0)

    assert np.allclose(got[:3], force(nxt), rtol=1e-6, atol=0)  # scale is formed in FP32
    assert not np.allclose(got[:3], force(own), rtol=1e-3, atol=0)
