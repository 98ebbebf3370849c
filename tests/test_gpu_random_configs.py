"""Seeded random small configurations of the periodic-shift step against the
oracle: one check over many combinations the fixed cases leave out.

Each case draws a shape of 2-20 x 1-12 x 1-12 cells (axes one or two cells
thick included), a kind per box face (given rho, given u, neither, both,
bounceback or plain fluid), a given-value slot per face, random given values,
5-15 % bounceback cells inside, a uniform or per-cell Smagorinsky constant or
none, and periodic copies on the axes without boundary faces.  Listed face
cells drop the normal component of every axis whose faces are not listed, as
the channel flags do; axes under three cells thick have no listed faces.

* FP32: 8 steps from a random non-equilibrium state against the FP32 oracle,
  element-wise within 1e-5 relative (the tolerance of the other step tests),
  and every local-update variant and bulk occupancy cap on the same bits.
* The same at 24-72 cells per axis (6 cases, 6 steps).
* FP16 / BF16 storage of f - w: the staged and fused variants on the same
  bits, and within 1e-3 (FP16) / 1e-2 (BF16) absolute of the FP64 oracle.
"""
import numpy as np
import pytest

from tests import random_cases as rc

pytestmark = pytest.mark.gpu

RTOL = 1e-5
CASES = 24


@pytest.mark.parametrize("seed", range(CASES))
def test_random_config_fp32_matches_oracle(psl, oracle_mod, seed):
    P, O = psl, oracle_mod
    c = rc.config(seed)
    print(f"[random fp32] seed {seed}: {rc.describe(c)}", flush=True)
    og = rc.oracle_grid(O, c, np.float32)
    f0 = og.logical_f().astype(np.float32)
    og.f[:] = f0
    steps = 8
    for _ in range(steps):
        og.step(periodic_x=c["copies"][0], periodic_y=c["copies"][1])
    ref = None
    for v in range(P.TUNE_VARIANTS):
        g = rc.grid(P, c, "f32")
        g.set_f(f0)
        g.tuning = v
        P.updateGrid([g], steps)
        f = g.get_f()
        assert g.shifter.tolist() == og.shift.tolist()
        g.close()
        if ref is None:
            assert np.all(np.isfinite(f))
            err = float(np.max(np.abs(f.astype(np.float64) - og.logical_f()) / np.maximum(np.abs(og.logical_f()), 1e-3)))
            print(f"  max rel err {err:.3e}", flush=True)
            assert err < RTOL
            ref = f
        else:
            assert np.array_equal(f.view(np.uint32), ref.view(np.uint32)), f"variant {v}"
    # the occupancy caps reserve LDS only: the same bits
    for occ in (2, 3, 4, 5, 6):
        g = rc.grid(P, c, "f32")
        g.set_f(f0)
        g.occupancy = occ
        assert g.occupancy == occ
        P.updateGrid([g], steps)
        assert np.array_equal(g.get_f().view(np.uint32), ref.view(np.uint32)), f"occupancy {occ}"
        g.close()


@pytest.mark.parametrize("seed", range(6))
def test_large_random_config_fp32_matches_oracle(psl, oracle_mod, seed):
    """24-72 cells per axis (up to 370 k cells): many workgroups per grid,
    population wrap thresholds inside waves every step, staged records over
    many waves; 6 steps against the oracle and every variant on the same bits."""
    P, O = psl, oracle_mod
    c = rc.config(seed, large=True)
    print(f"[large random fp32] seed {seed}: {rc.describe(c)}", flush=True)
    og = rc.oracle_grid(O, c, np.float32)
    f0 = og.logical_f().astype(np.float32)
    og.f[:] = f0
    steps = 6
    for _ in range(steps):
        og.step(periodic_x=c["copies"][0], periodic_y=c["copies"][1])
    ref = None
    for v in range(P.TUNE_VARIANTS):
        g = rc.grid(P, c, "f32")
        g.set_f(f0)
        g.tuning = v
        P.updateGrid([g], steps)
        f = g.get_f()
        assert g.shifter.tolist() == og.shift.tolist()
        g.close()
        if ref is None:
            err = float(np.max(np.abs(f.astype(np.float64) - og.logical_f()) / np.maximum(np.abs(og.logical_f()), 1e-3)))
            print(f"  max rel err {err:.3e}", flush=True)
            assert err < RTOL
            ref = f
        else:
            assert np.array_equal(f.view(np.uint32), ref.view(np.uint32)), f"variant {v}"


@pytest.mark.parametrize("storage,bound", [("f16", 1e-3), ("bf16", 1e-2)])
def test_random_configs_reduced_storage(psl, oracle_mod, storage, bound):
    P, O = psl, oracle_mod
    worst = 0.0
    for seed in range(0, CASES, 3):
        c = rc.config(seed)
        print(f"[random {storage}] seed {seed}: {rc.describe(c)}", flush=True)
        og = rc.oracle_grid(O, c, np.float64)
        f0 = og.logical_f().astype(np.float32)
        og.f[:] = f0
        steps = 8
        for _ in range(steps):
            og.step(periodic_x=c["copies"][0], periodic_y=c["copies"][1])
        outs = []
        for v in (0, 8):  # staged and fused box faces
            g = rc.grid(P, c, storage)
            g.set_f(f0)
            g.tuning = v
            P.updateGrid([g], steps)
            outs.append(g.get_f())
            g.close()
        assert np.array_equal(outs[0].view(np.uint32), outs[1].view(np.uint32))
        err = float(np.max(np.abs(outs[0].astype(np.float64) - og.logical_f())))
        worst = max(worst, err)
        assert err < bound, f"seed {seed}: {err:.3e}"
    print(f"  worst |f - f64| {worst:.3e}", flush=True)
