"""Restart state and health checks (psl_grid_*_state, psl_diad_*_state,
psl_*_check_health; no reference counterpart: the reference never writes
fArray and has no guard against a diverging run).

* Periodic-shift sphere channel in FP32, FP16 and BF16: 9 steps, the state
  saved, 7 more steps; the state loaded back into the same grid and into a fresh
  grid with the same settings, and 7 steps again: populations and shifter
  bit-identical.  The same with a learned closure evaluated every third step
  (the state taken inside a hold period) and with scalar transport.
* Esotwist: three wall-refined DIAD levels through the recursion, the state of
  every level saved, two coarse steps, the states loaded, two coarse steps
  again: bit-identical (the loaded levels' interface records are stale until
  their next update, so the first transfers read slots).
* Health: a fresh grid reports no bad cell, and its density range equals
  numpy's from get_f; a NaN planted in one population and a cell whose
  populations sum below zero are each counted once, and the range leaves them
  out.
* Loading refuses a state of another storage, size or kind, and a truncated one.
"""
import numpy as np
import pytest

from periodicshiftlbm_amd import cases, les

pytestmark = pytest.mark.gpu


def _channel(P, storage, cs=0.1):
    d = cases.batch_domains(count=1, seed=5, l_min=16, l_max=16)[0]
    g, _ = cases.domain_grid(d, storage, cs=cs)
    return g


def _assert_same(a, b, label):
    fa, fb = a.get_f(), b.get_f()
    nbad = int(np.count_nonzero(fa.view(np.uint32) != fb.view(np.uint32)))
    print(f"  {label}: {nbad} differing values", flush=True)
    assert a.shifter.tolist() == b.shifter.tolist()
    assert nbad == 0


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_periodic_shift_restart_is_bit_exact(psl, storage):
    P = psl
    print(f"[state] periodic shift {storage}", flush=True)
    g = _channel(P, storage)
    P.updateGrid([g], 9)
    s = g.state()
    assert s[:8].tobytes() == b"PSLSTATE"
    P.updateGrid([g], 7)
    ref = _channel(P, storage)
    ref.load_state(g.state())  # the state after 16 steps, as a reference copy
    g.load_state(s)
    fresh = _channel(P, storage)
    fresh.load_state(s)
    for x in (g, fresh):
        P.updateGrid([x], 7)
    _assert_same(g, ref, "reloaded grid")
    _assert_same(fresh, ref, "fresh grid from the state")
    for x in (g, ref, fresh):
        x.close()


def test_restart_inside_a_closure_hold_period(psl):
    P = psl
    n = 24
    ux = 0.04 * np.sin(2 * np.pi * np.arange(n) / n)
    u = [np.broadcast_to(ux[None, None, :], (n, n, n)).astype(np.float32), np.zeros((n, n, n), np.float32),
         np.zeros((n, n, n), np.float32)]

    def grid():
        g = les.periodic_grid(P, n, 1e-3, "f16", u, 0.0)
        g.set_closure(net)
        g.set_closure_every(3)
        return g

    net = P.MLP(les.glorot_params(7))
    g = grid()
    P.updateGrid([g], 4)  # evaluated at update 0 and 3: the state sits one update into a hold
    s = g.state()
    P.updateGrid([g], 5)
    ref = grid()
    ref.load_state(g.state())
    g.load_state(s)
    fresh = grid()
    fresh.load_state(s)
    assert np.array_equal(fresh.closure_y(), g.closure_y())
    for x in (g, fresh):
        P.updateGrid([x], 5)
    _assert_same(g, ref, "closure, reloaded")
    _assert_same(fresh, ref, "closure, fresh grid")
    # a grid evaluating at another interval refuses the state
    other = les.periodic_grid(P, n, 1e-3, "f16", u, 0.0)
    other.set_closure(net)
    other.set_closure_every(4)
    with pytest.raises(RuntimeError, match="interval"):
        other.load_state(s)
    for x in (g, ref, fresh, other):
        x.close()


def test_restart_with_scalar_transport(psl):
    P = psl

    def grid():
        g = _channel(P, "f32")
        st = P.ScalarTransportStruct(g, 0.7)
        given = np.full(g.n, np.nan, np.float32)
        given[: g.Info.cellCountX * 4] = 1.0
        st.set_given(given)
        return g, st

    def run(g, st, k):
        for _ in range(k):
            P.applyStreaming(g)
            P.applyScalarTransportLocalCellUpdate(g, st)

    g, st = grid()
    run(g, st, 6)
    s = g.state()
    run(g, st, 4)
    ref, rst = grid()
    ref.load_state(g.state())
    g.load_state(s)
    run(g, st, 4)
    _assert_same(g, ref, "scalar, f")
    assert np.array_equal(st.get_T().view(np.uint32), rst.get_T().view(np.uint32))
    plain = _channel(P, "f32")
    with pytest.raises(RuntimeError, match="scalar"):
        plain.load_state(s)
    for x in (g, ref, plain):
        x.close()


def test_diad_levels_restart_is_bit_exact(psl):
    P = psl
    lv, grids = cases.diad_sphere(P, "f16", n0=16, levels=3, span=1, radius=2.0)
    for g in grids:
        g.fill_equilibrium(ux=0.04)
    P.updateGrid(grids, 2)
    states = [g.state() for g in grids]
    P.updateGrid(grids, 2)
    after = [g.get_f() for g in grids]
    flips = [g.esotwistFlipper for g in grids]
    for g, s in zip(grids, states):
        g.load_state(s)
        assert g.interface_records()[1] is False
    P.updateGrid(grids, 2)
    for l, (g, a) in enumerate(zip(grids, after)):
        nbad = int(np.count_nonzero(g.get_f().view(np.uint32) != a.view(np.uint32)))
        print(f"  DIAD level {l}: {nbad} differing values", flush=True)
        assert g.esotwistFlipper == flips[l]
        assert nbad == 0


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_health_counts_bad_cells(psl, storage):
    P = psl
    g = _channel(P, storage)
    P.updateGrid([g], 5)
    f = g.get_f()
    rho = f.astype(np.float64).sum(axis=0)
    bad, lo, hi = g.check_health()
    print(f"[health] {storage}: bad {bad}, rho [{lo:.6f}, {hi:.6f}] numpy [{rho.min():.6f}, {rho.max():.6f}]",
          flush=True)
    assert bad == 0
    tol = 1e-6 if storage == "f32" else 2e-3
    assert abs(lo - rho.min()) <= tol and abs(hi - rho.max()) <= tol
    # one NaN population and one cell with negative populations
    f[5, 100] = np.nan
    f[:, 200] = -0.01
    g.set_f(f)
    bad, lo, hi = g.check_health()
    keep = np.ones(g.n, bool)
    keep[[100, 200]] = False
    print(f"  planted: bad {bad}, rho [{lo:.6f}, {hi:.6f}]", flush=True)
    assert bad == 2
    assert abs(lo - rho[keep].min()) <= tol and abs(hi - rho[keep].max()) <= tol
    g.close()
    # the same check on an esotwist grid
    lv, grids = cases.diad_sphere(P, storage, n0=12, levels=2, span=1, radius=2.0)
    d = grids[1]
    assert d.check_health()[0] == 0
    fd = d.get_f()
    fd[3, 7] = np.inf
    d.set_f(fd)
    assert d.check_health()[0] == 1


def test_load_state_refuses_mismatches(psl):
    P = psl
    g32 = _channel(P, "f32")
    g16 = _channel(P, "f16")
    s32 = g32.state()
    with pytest.raises(RuntimeError, match="storage"):
        g16.load_state(s32)
    with pytest.raises(RuntimeError, match="size"):
        g32.load_state(s32[:-1])
    d = cases.batch_domains(count=1, seed=5, l_min=20, l_max=20)[0]
    big, _ = cases.domain_grid(d, "f32")
    with pytest.raises(RuntimeError, match="size"):
        big.load_state(s32)
    lv, grids = cases.diad_sphere(P, "f32", n0=12, levels=2, span=1, radius=2.0)
    with pytest.raises(RuntimeError, match="esotwist|periodic-shift"):
        grids[0].load_state(s32)
    bad = s32.copy()
    bad[0] = ord("X")
    with pytest.raises(RuntimeError, match="magic"):
        g32.load_state(bad)
    for x in (g32, g16, big):
        x.close()
