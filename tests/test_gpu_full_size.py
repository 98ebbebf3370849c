"""The benchmark configurations at their full sizes, through properties that
hold at any size (the scaled-down configurations are held to the FP64 oracle
in tests/test_gpu_configs.py; the 64^3 Taylor-Green box runs there at full
size).

* Sphere channel, 256 x 128 x 128 (icosphere D = 32 voxelized by the STL path,
  velocity inlet U = 0.04, pressure outlet, periodic y copies, Re = 1000,
  Cs = 0.17), FP32, from uniform flow around the sphere:
    - 20 steps against the FP32 oracle at full size, 1e-5 relative (the
      oracle's mixed FP32/FP64 expressions), and the momentum-exchange force
      on the sphere box against the oracle's;
    - every local-update variant (staged or fused box faces, non-temporal
      loads / stores, block order) ends on the same bits after 20 steps.
* The same channel from rest over 20000 steps (flowAroundSphereMainSingle.cu
  records C_D every step): with f - w stored in FP16 the mean drag
  coefficient over the second half is within 1 % of the FP32 run's.
* Two-level sphere at Re = 1e5: coarse 256 x 128 x 128 with the 2x finer box
  over coarse cells [32, 160) x [16, 112) x [16, 112) (256 x 192 x 192 fine
  cells), FP32 through the recursion (coarse step, two fine steps, the
  coarse-fine communication) against the oracle's recursion for 3 coarse steps,
  1e-5 relative on both levels.
* Batch: the 64 domains of the batch configuration (L ~ U[32, 96], FP16),
  advanced together for 20 steps at the autotuned configuration, equal bit for
  bit to the same domains stepped one by one.
"""
import numpy as np
import pytest

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b), 1e-3)))


def _sphere_channel(P, storage, shape=cases.SPHERE_SHAPE, reynolds=1000.0, cs=0.17):
    nx, ny, nz = shape
    nu = cases.sphere_nu(0.04, cases.SPHERE_DIAMETER, reynolds)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=nu, res=1.0)
    stl = P.STLStruct.from_triangles(cases.icosphere(4, cases.SPHERE_DIAMETER / 2, cases.SPHERE_CENTER))
    mask = P.BoolArray(nx * ny * nz)
    P.applyMarkersInsideSTL(mask, stl, True, info)
    solid = mask.get().astype(bool)
    flags = cases.channel_flags(shape, solid)
    g = P.GridStruct(info, storage)
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(0.04))
    g.set_smagorinsky(cs)
    g.set_periodic_copies(y=True)
    return g, flags, solid, nu


def test_sphere_channel_full_size_matches_oracle(psl, oracle_mod):
    P, O = psl, oracle_mod
    nx, ny, nz = cases.SPHERE_SHAPE
    g, flags, solid, nu = _sphere_channel(P, "f32")
    n = g.n
    print(f"[full-size sphere] {cases.SPHERE_SHAPE}, {int(solid.sum())} solid cells, 20 steps vs the oracle",
          flush=True)
    og = O.Grid(nx, ny, nz, dtype=np.float32, nu=nu, cs=0.17, flags=flags, bcvals=cases.channel_given_values(0.04))
    ux = np.where(solid, 0.0, 0.04)
    og.fill_equilibrium(np.ones(n), ux, np.zeros(n), np.zeros(n))
    g.set_f(og.logical_f().astype(np.float32))
    steps = 20
    P.updateGrid([g], steps)
    for _ in range(steps):
        og.step(periodic_y=True)
    err = _rel_err(g.get_f(), og.logical_f())
    lo, hi = cases.sphere_box()
    fg, fo = g.momentum_exchange(lo, hi), og.momentum_exchange(lo, hi)
    print(f"  max rel err {err:.3e}; force gpu {fg} oracle {fo}", flush=True)
    assert err < RTOL
    assert np.all(np.abs(fg - fo) <= 1e-5 * np.max(np.abs(fo)))
    g.close()


def test_sphere_channel_full_size_variants_identical(psl):
    P = psl
    ref = None
    for v in range(P.TUNE_VARIANTS):
        g, _, solid, _ = _sphere_channel(P, "f32")
        P.fill_equilibrium(g, ux=np.where(solid, 0.0, 0.04).astype(np.float32))
        g.tuning = v
        P.updateGrid([g], 20)
        f = g.get_f().view(np.uint32)
        g.close()
        if ref is None:
            ref = f
            continue
        nbad = int(np.count_nonzero(f != ref))
        print(f"  variant {v}: {nbad} differing values", flush=True)
        assert nbad == 0


def test_sphere_drag_history_fp16_within_one_percent_of_fp32(psl):
    P = psl
    lo, hi = cases.sphere_box()
    steps, every = 20000, 20
    means = {}
    for storage in ("f32", "f16"):
        g, _, _, _ = _sphere_channel(P, storage)
        P.fill_equilibrium(g)
        fx = []
        for _ in range(steps // every):
            P.updateGrid([g], every)
            fx.append(float(g.momentum_exchange(lo, hi)[0]))
        g.close()
        cd = np.array([cases.drag_coefficient(f) for f in fx])
        assert np.all(np.isfinite(cd))
        means[storage] = float(np.mean(cd[len(cd) // 2:]))
        print(f"[drag history] {storage}: mean C_D over steps {steps // 2}-{steps}: {means[storage]:.5f}",
              flush=True)
    rel = means["f16"] / means["f32"] - 1
    print(f"  FP16 vs FP32: {rel:+.3%}", flush=True)
    assert abs(rel) < 0.01


def test_two_level_sphere_full_size_matches_oracle(psl, oracle_mod):
    P, O = psl, oracle_mod
    (gc, gf), _ = cases.two_level_sphere("f32", reynolds=1e5)
    print(f"[full-size two-level] coarse {cases.SPHERE_SHAPE} fine "
          f"{(gf.Info.cellCountX, gf.Info.cellCountY, gf.Info.cellCountZ)}, 3 coarse steps vs the oracle", flush=True)
    bc = cases.channel_given_values(0.04)
    oc = O.Grid(gc.Info.cellCountX, gc.Info.cellCountY, gc.Info.cellCountZ, dtype=np.float32, nu=gc.Info.nu,
                cs=0.17, flags=gc.get_flags(), bcvals=bc, sub=cases.TWO_LEVEL_SUB)
    of = O.Grid(gf.Info.cellCountX, gf.Info.cellCountY, gf.Info.cellCountZ, dtype=np.float32, nu=gf.Info.nu,
                cs=0.17, flags=gf.get_flags(), bcvals=bc)
    # uniform flow on both levels so that the interface carries a non-trivial state
    for g, o in ((gc, oc), (gf, of)):
        n = g.n
        fl = g.get_flags()
        ux = np.where(fl == cases.BOUNCEBACK, 0.0, 0.04)
        o.fill_equilibrium(np.ones(n), ux, np.zeros(n), np.zeros(n))
        g.set_f(o.logical_f().astype(np.float32))
    steps = 3
    P.updateGrid([gc, gf], steps)
    for _ in range(steps):
        oc.step(periodic_y=True)
        of.step()
        of.step()
        O.coarse_fine_comm(oc, of)
    for name, g, o in (("coarse", gc, oc), ("fine", gf, of)):
        err = _rel_err(g.get_f(), o.logical_f())
        print(f"  {name}: max rel err {err:.3e}", flush=True)
        assert err < RTOL
    gc.close()
    gf.close()


def test_full_batch_equals_single_grid_steps(psl):
    P = psl
    doms = cases.batch_domains()
    batched = [cases.domain_grid(d, "f16")[0] for d in doms]
    b = P.Batch(batched)
    best, _ = b.autotune_layouts(steps=1)
    cells, _, _ = b.census()
    tuned = 4 * 2 * 16 * 2  # steps the search advanced every grid
    print(f"[full batch] {len(doms)} domains, {cells} cells, best configuration {best}", flush=True)
    b.step(20)
    b.close()
    worst = 0
    for d, gb in zip(doms, batched):
        gs = cases.domain_grid(d, "f16")[0]
        P.updateGrid([gs], tuned + 20)
        nbad = int(np.count_nonzero(gb.get_f().view(np.uint32) != gs.get_f().view(np.uint32)))
        assert gb.shifter.tolist() == gs.shifter.tolist()
        worst = max(worst, nbad)
        gs.close()
        gb.close()
    print(f"  most differing values in a domain: {worst}", flush=True)
    assert worst == 0
