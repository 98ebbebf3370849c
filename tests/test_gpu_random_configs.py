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
  and every local-update variant on the same bits.
* FP16 / BF16 storage of f - w: the staged and fused variants on the same
  bits, and within 1e-3 (FP16) / 1e-2 (BF16) absolute of the FP64 oracle.
"""
import numpy as np
import pytest

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5
CASES = 24
FACE_KINDS = [cases.GIVEN_U, cases.GIVEN_RHO, cases.GIVEN_NONE, cases.GIVEN_RHOU, cases.BOUNCEBACK, cases.FLUID]


def _config(seed):
    rng = np.random.default_rng(1000 + seed)
    shape = (int(rng.integers(2, 21)), int(rng.integers(1, 13)), int(rng.integers(1, 13)))
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    coord = [i.reshape(-1), j.reshape(-1), k.reshape(-1)]
    n = nx * ny * nz
    flags = np.full(n, cases.FLUID, np.uint8)
    solid = rng.random(n) < rng.uniform(0.05, 0.15)
    flags[solid] = cases.BOUNCEBACK
    # a kind and a given-value slot per face; thin axes and axes drawn as
    # open get plain fluid faces
    face = {}
    for a in range(3):
        for side in (0, 1):
            kind = cases.FLUID if shape[a] < 3 else int(rng.choice(FACE_KINDS))
            face[(a, side)] = (kind, int(rng.integers(0, 4)))
    listed_axes = [a for a in range(3) if any(face[(a, s)][0] >= cases.GIVEN_RHO for s in (0, 1))]
    periodic_bits = 0
    for a, bit in enumerate((cases.PERIODIC_X, cases.PERIODIC_Y, cases.PERIODIC_Z)):
        if a not in listed_axes:
            periodic_bits |= bit
    for a in range(3):
        for side in (0, 1):
            kind, slot = face[(a, side)]
            if kind == cases.FLUID:
                continue
            on = coord[a] == (0 if side == 0 else shape[a] - 1)
            if kind == cases.BOUNCEBACK:
                flags[on] = cases.BOUNCEBACK
            else:
                flags[on] = kind | periodic_bits | (slot << 6)
    bc = np.column_stack([rng.uniform(0.98, 1.02, 4), rng.uniform(-0.03, 0.03, (4, 3))]).astype(np.float32)
    cs_mode = int(rng.integers(0, 3))
    cs = 0.0 if cs_mode == 0 else float(rng.uniform(0.05, 0.2))
    cs_cell = rng.uniform(0.05, 0.2, n).astype(np.float32) if cs_mode == 2 else None
    copies = (0 not in listed_axes and nx >= 3 and bool(rng.integers(0, 2)),
              1 not in listed_axes and ny >= 3 and bool(rng.integers(0, 2)))
    rho = rng.uniform(0.98, 1.02, n)
    u = rng.uniform(-0.03, 0.03, (3, n))
    noise = rng.normal(0, 0.005, (27, n))
    return dict(shape=shape, flags=flags, bc=bc, cs=cs, cs_cell=cs_cell, copies=copies, rho=rho, u=u, noise=noise,
                nu=float(rng.uniform(0.005, 0.05)))


def _grid(P, c, storage):
    nx, ny, nz = c["shape"]
    g = P.GridStruct(P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=c["nu"]), storage)
    g.set_flags(c["flags"])
    g.set_given_values(c["bc"])
    g.set_smagorinsky(c["cs"] if c["cs_cell"] is None else c["cs_cell"])
    g.set_periodic_copies(x=c["copies"][0], y=c["copies"][1])
    return g


def _oracle(O, c, dtype):
    nx, ny, nz = c["shape"]
    og = O.Grid(nx, ny, nz, dtype=dtype, nu=c["nu"], cs=c["cs"], flags=c["flags"], bcvals=c["bc"],
                cs_cell=c["cs_cell"])
    og.fill_equilibrium(c["rho"], *c["u"])
    og.f[:] = og.logical_f() * (1 + c["noise"])
    return og


def _describe(c):
    kinds = np.bincount(c["flags"] & 7, minlength=8)
    return (f"shape={c['shape']} fluid={kinds[0]} bounceback={kinds[1]} listed={int(kinds[2:6].sum())} "
            f"cs={'per-cell' if c['cs_cell'] is not None else c['cs']:} copies={c['copies']}")


@pytest.mark.parametrize("seed", range(CASES))
def test_random_config_fp32_matches_oracle(psl, oracle_mod, seed):
    P, O = psl, oracle_mod
    c = _config(seed)
    print(f"[random fp32] seed {seed}: {_describe(c)}", flush=True)
    og = _oracle(O, c, np.float32)
    f0 = og.logical_f().astype(np.float32)
    og.f[:] = f0
    steps = 8
    for _ in range(steps):
        og.step(periodic_x=c["copies"][0], periodic_y=c["copies"][1])
    ref = None
    for v in range(P.TUNE_VARIANTS):
        g = _grid(P, c, "f32")
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


@pytest.mark.parametrize("storage,bound", [("f16", 1e-3), ("bf16", 1e-2)])
def test_random_configs_reduced_storage(psl, oracle_mod, storage, bound):
    P, O = psl, oracle_mod
    worst = 0.0
    for seed in range(0, CASES, 3):
        c = _config(seed)
        print(f"[random {storage}] seed {seed}: {_describe(c)}", flush=True)
        og = _oracle(O, c, np.float64)
        f0 = og.logical_f().astype(np.float32)
        og.f[:] = f0
        steps = 8
        for _ in range(steps):
            og.step(periodic_x=c["copies"][0], periodic_y=c["copies"][1])
        outs = []
        for v in (0, 8):  # staged and fused box faces
            g = _grid(P, c, storage)
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
