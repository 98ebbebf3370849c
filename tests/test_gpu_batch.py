"""Batched steps over independent grids of uneven size (psl_batch_*).

A batch advances every grid by one hipGraph replay per step, with the same
per-cell device code as the single-grid kernels.  Each grid must end exactly
(bit for bit, every storage type, every kernel variant, block order, number
of cells per thread and tile order) where
the same number of single-grid updateGrid steps puts it; the single-grid
path is pinned against the oracle by the other GPU tests, and one batched
domain is compared with the oracle directly here (1e-5 relative, FP32).
Domains are sphere-in-channel boxes 2L x L x L (velocity inlet, pressure
outlet, bounceback walls in y, z periodic through the shift), one of them
without Smagorinsky (Cs = 0).
"""
import numpy as np
import pytest

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b), 1e-3)))


def _domains():
    return cases.batch_domains(count=5, seed=11, l_min=12, l_max=28)


def _grids(doms, storage):
    # the third domain runs without Smagorinsky
    return [cases.domain_grid(d, storage, cs=0.0 if i == 2 else 0.17)[0] for i, d in enumerate(doms)]


def _assert_identical(doms, batched, single, label):
    for d, gb, gs in zip(doms, batched, single):
        assert gb.shifter.tolist() == gs.shifter.tolist()
        fb, fs = gb.get_f(), gs.get_f()
        nbad = int(np.count_nonzero(fb != fs))
        print(f"  {label} L={d['L']} re={d['re']:.0f}: {nbad} differing values, max rel "
              f"{_rel_err(fb, fs):.3e}", flush=True)
        assert nbad == 0


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_batch_matches_single_grid_steps(psl, storage):
    P = psl
    doms = _domains()
    print(f"[batch] storage={storage} shapes={[d['shape'] for d in doms]}", flush=True)
    batched = _grids(doms, storage)
    single = _grids(doms, storage)
    b = P.Batch(batched)
    cells, blocks, nbnd = b.census()
    assert cells == sum(int(np.prod(d["shape"])) for d in doms)
    steps = 7
    b.step(steps)
    for g in single:
        P.updateGrid([g], steps)
    _assert_identical(doms, batched, single, "two-launch")
    # other variants (fused boundary, non-temporal stores, XCD order) and the
    # largest-first block order give the same result
    for tuning, order in ((P.TUNE_FUSED_BOUNDARY | P.TUNE_XCD_REMAP, 0),
                          (P.TUNE_NT_STORE | P.TUNE_FUSED_BOUNDARY, 1),
                          (P.TUNE_NT_LOAD | P.TUNE_XCD_REMAP, 1)):
        b.set_tuning(tuning, order)
        b.step(3)
        for g in single:
            P.updateGrid([g], 3)
        _assert_identical(doms, batched, single, f"tuning {tuning} order {order}")
    # the layouts of the two-launch bulk kernel: two cells per thread and the
    # tiles interleaved across grids
    for cpt, tile in ((2, 0), (1, 1), (2, 1)):
        for tuning in (0, P.TUNE_XCD_REMAP | P.TUNE_NT_LOAD):
            b.set_tuning(tuning, 1)
            b.set_layout(cpt, tile)
            assert b.layout == (cpt, tile)
            b.step(3)
            for g in single:
                P.updateGrid([g], 3)
            _assert_identical(doms, batched, single, f"layout cpt {cpt} tiles {tile} tuning {tuning}")
    # the full search: 4 layouts x 2 orders x 16 variants, one warm-up step
    # and one timed step each (the grids advance), fused variants at layout 1/0 only
    best, times = b.autotune_layouts(steps=1)
    assert best[2] in (1, 2) and best[3] in (0, 1)
    assert int(np.count_nonzero(times >= 0)) == 8 * 16 - 3 * 2 * 8
    for g in single:
        P.updateGrid([g], 4 * 2 * 16 * 2)
    _assert_identical(doms, batched, single, f"after the layout search (best {best})")
    b.step(5)
    for g in single:
        P.updateGrid([g], 5)
    _assert_identical(doms, batched, single, f"best layout {best}")
    b.close()


def test_batch_domain_matches_oracle(psl, oracle_mod):
    P, O = psl, oracle_mod
    d = _domains()[1]
    g, flags = cases.domain_grid(d, "f32")
    nx, ny, nz = d["shape"]
    og = O.Grid(nx, ny, nz, dtype=np.float32, nu=d["nu"], cs=0.17, flags=flags,
                bcvals=cases.channel_given_values(d["u_inlet"]))
    og.fill_equilibrium()
    b = P.Batch([g])
    steps = 12
    print(f"[batch-oracle] shape={d['shape']} steps={steps}", flush=True)
    b.step(steps)
    for _ in range(steps):
        og.step()
    err = _rel_err(g.get_f(), og.logical_f())
    print(f"  max rel err {err:.3e}")
    assert err < RTOL
    b.close()


def test_batch_autotune_keeps_results(psl):
    P = psl
    doms = _domains()[:3]
    batched = _grids(doms, "f16")
    single = _grids(doms, "f16")
    b = P.Batch(batched)
    tuning, order, times = b.autotune(2)
    print(f"[batch-autotune] best tuning {tuning} order {order} ms/step {np.round(times, 4).tolist()}", flush=True)
    assert np.all(times > 0)
    # autotune advanced the batch by 2 orders x 16 variants x (1 + 2) steps
    for g in single:
        P.updateGrid([g], 2 * P.TUNE_VARIANTS * 3)
    _assert_identical(doms, batched, single, "after autotune")
    b.close()


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_single_grid_variants_identical(psl, storage):
    """Every local-update variant of the single-grid path (staged or fused box
    faces, plain or non-temporal accesses, either block order) gives the same bits."""
    P = psl
    d = _domains()[1]
    ref = cases.domain_grid(d, storage)[0]
    ref.tuning = 0
    P.updateGrid([ref], 9)
    f0 = ref.get_f()
    for v in range(1, P.TUNE_VARIANTS):
        g = cases.domain_grid(d, storage)[0]
        g.tuning = v
        P.updateGrid([g], 9)
        nbad = int(np.count_nonzero(g.get_f() != f0))
        print(f"  variant {v}: {nbad} differing values", flush=True)
        assert nbad == 0


def test_batch_rejects_mixed_storage(psl):
    P = psl
    doms = _domains()[:2]
    a = cases.domain_grid(doms[0], "f32")[0]
    c = cases.domain_grid(doms[1], "f16")[0]
    with pytest.raises(RuntimeError):
        P.Batch([a, c])


def test_batch_multi_step_graphs_and_per_cell_cs(psl):
    """Graphs of several captured steps (steps_per_graph) advance every grid as
    the same number of single steps, and a grid with a per-cell Smagorinsky
    field (getSmagorinskyConstant per cell) keeps it inside the batch."""
    P = psl
    doms = _domains()
    rng = np.random.default_rng(5)
    cs_fields = {1: rng.uniform(0.0, 0.3, int(np.prod(doms[1]["shape"]))).astype(np.float32)}
    batched, single = _grids(doms, "f16"), _grids(doms, "f16")
    for i, cs in cs_fields.items():
        batched[i].set_smagorinsky(cs)
        single[i].set_smagorinsky(cs)
    b = P.Batch(batched)
    assert b.steps_per_graph == 1000
    b.steps_per_graph = 7
    print("[batch multi-step graphs] 7 steps per graph, 17 steps", flush=True)
    b.step(17)  # two 7-step graphs and three single steps
    for g in single:
        P.updateGrid([g], 17)
    _assert_identical(doms, batched, single, "7-step graphs")
    b.steps_per_graph = 1
    b.step(4)
    for g in single:
        P.updateGrid([g], 4)
    _assert_identical(doms, batched, single, "one-step graphs")
    # the per-cell field matters: the same grid with uniform Cs ends elsewhere
    alt = _grids(doms[1:2], "f16")[0]
    P.updateGrid([alt], 21)
    assert np.count_nonzero(alt.get_f() != batched[1].get_f()) > 0
    with pytest.raises(P.PslError):
        b.steps_per_graph = 0
    b.close()


def test_batch_refuses_grid_changed_to_unbatchable(psl):
    """Periodic copies set on a grid after the batch was created are refused
    at the next step instead of being ignored by the captured graph."""
    P = psl
    doms = _domains()[:2]
    grids = _grids(doms, "f32")
    b = P.Batch(grids)
    b.step(1)
    grids[0].set_periodic_copies(y=True)
    with pytest.raises(P.PslError):
        b.step(1)
    grids[0].set_periodic_copies()
    b.step(1)
    b.close()
