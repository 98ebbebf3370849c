"""Seeded random two-level periodic-shift configurations against the oracle's
recursion (flowAroundSphereMainSingle.cu:158-167 over
include/gridRefinementFunctions.h:283-536).

Each case draws a coarse grid of 6-24 x 5-16 x 5-16 cells and a sub-grid of
2 cells up to the whole extent per axis at a random position, flush with the
grid edges included (the smallest sub-grids the transfers accept: their
fine-to-coarse ring faces then meet or cross).  The fine grid is the sub-grid
at twice the resolution with a ghost shell one fine cell thick, as the
reference driver builds it.  Random viscosity, Smagorinsky constant on or off,
periodic copies on the coarse grid on or off, and a smooth random flow plus a
small non-equilibrium part on both levels.

* FP32: three coarse steps of updateGrid (coarse step, two fine steps, the
  coarse-fine communication) against the oracle, within 1e-5 relative on both
  levels (the tolerance of the other refinement tests).
* fillCoarseGridFromFine on the same pair against the oracle's.
"""
import numpy as np
import pytest

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5
CASES = 12


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b), 1e-3)))


def _case(seed):
    rng = np.random.default_rng(9000 + seed)
    shape = (int(rng.integers(6, 25)), int(rng.integers(5, 17)), int(rng.integers(5, 17)))
    ext = [int(rng.integers(2, n + 1)) for n in shape]
    start = [int(rng.integers(0, n - e + 1)) for n, e in zip(shape, ext)]
    sub = (start[0], start[1], start[2], start[0] + ext[0], start[1] + ext[1], start[2] + ext[2])
    return dict(shape=shape, sub=sub, nu=float(rng.uniform(0.005, 0.03)),
                cs=float(rng.choice([0.0, rng.uniform(0.05, 0.17)])), copies=bool(rng.integers(0, 2)),
                amp=float(rng.uniform(0.01, 0.04)), phase=rng.uniform(0, 2 * np.pi, 3), seed=seed)


def _field(c, nx, ny, nz, x0, y0, z0, h, rng):
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    lx, ly, lz = c["shape"]
    x = 2 * np.pi * (x0 + i * h) / lx + c["phase"][0]
    y = 2 * np.pi * (y0 + j * h) / ly + c["phase"][1]
    z = 2 * np.pi * (z0 + k * h) / lz + c["phase"][2]
    a = c["amp"]
    ux = a * np.sin(x) * np.cos(y) * np.cos(z)
    uy = -a * np.cos(x) * np.sin(y) * np.cos(z)
    uz = 0.5 * a * np.sin(z) * np.cos(x)
    n = nx * ny * nz
    return 1 + 0.01 * rng.standard_normal(n) * 0.1, ux.reshape(-1), uy.reshape(-1), uz.reshape(-1)


def _pair(P, O, c):
    nxc, nyc, nzc = c["shape"]
    sub = c["sub"]
    nxf, nyf, nzf = 2 * (sub[3] - sub[0]), 2 * (sub[4] - sub[1]), 2 * (sub[5] - sub[2])
    ci = P.InfoStruct(cellCountX=nxc, cellCountY=nyc, cellCountZ=nzc, nu=c["nu"], res=1.0,
                      iSubgridStart=sub[0], jSubgridStart=sub[1], kSubgridStart=sub[2],
                      iSubgridEnd=sub[3], jSubgridEnd=sub[4], kSubgridEnd=sub[5])
    nu_f = 2 * c["nu"]
    fi = P.InfoStruct(gridID=1, cellCountX=nxf, cellCountY=nyf, cellCountZ=nzf, nu=nu_f, res=0.5,
                      ox=sub[0] - 0.25, oy=sub[1] - 0.25, oz=sub[2] - 0.25)
    gc, gf = P.GridStruct(ci, "f32"), P.GridStruct(fi, "f32")
    k, j, i = np.meshgrid(np.arange(nzf), np.arange(nyf), np.arange(nxf), indexing="ij")
    ghost = (i == 0) | (i == nxf - 1) | (j == 0) | (j == nyf - 1) | (k == 0) | (k == nzf - 1)
    fflags = np.where(ghost, cases.GHOST, cases.FLUID).astype(np.uint8).reshape(-1)
    gf.set_flags(fflags)
    gc.set_periodic_copies(x=c["copies"], y=c["copies"])
    gc.set_smagorinsky(c["cs"])
    gf.set_smagorinsky(c["cs"])
    oc = O.Grid(nxc, nyc, nzc, dtype=np.float32, nu=c["nu"], cs=c["cs"], sub=sub)
    of = O.Grid(nxf, nyf, nzf, dtype=np.float32, nu=nu_f, cs=c["cs"], flags=fflags)
    rng = np.random.default_rng(100 + c["seed"])
    for g, o, args in ((gc, oc, (nxc, nyc, nzc, 0, 0, 0, 1.0)),
                       (gf, of, (nxf, nyf, nzf, sub[0] - 0.25, sub[1] - 0.25, sub[2] - 0.25, 0.5))):
        o.fill_equilibrium(*_field(c, *args, rng))
        f0 = (o.logical_f() * (1 + rng.normal(0, 0.003, o.f.shape))).astype(np.float32)
        o.f[:] = f0  # shift 0: the logical order is the storage order
        g.set_f(f0)
    return gc, gf, oc, of


def _describe(c):
    return f"coarse={c['shape']} sub={c['sub']} cs={c['cs']:.3f} copies={c['copies']}"


@pytest.mark.parametrize("seed", range(CASES))
def test_random_two_level_recursion_matches_oracle(psl, oracle_mod, seed):
    P, O = psl, oracle_mod
    c = _case(seed)
    print(f"[random refinement] seed {seed}: {_describe(c)}", flush=True)
    gc, gf, oc, of = _pair(P, O, c)
    steps = 3
    P.updateGrid([gc, gf], steps)
    for _ in range(steps):
        oc.step(periodic_x=c["copies"], periodic_y=c["copies"])
        for _ in range(2):
            of.step()
        O.coarse_fine_comm(oc, of)
    ec, ef = _rel_err(gc.get_f(), oc.logical_f()), _rel_err(gf.get_f(), of.logical_f())
    print(f"  coarse {ec:.3e} fine {ef:.3e}", flush=True)
    assert ec < RTOL and ef < RTOL
    # the coarse fill from the fine grid on the state reached
    P.fillCoarseGridFromFine(gc, gf)
    O.fill_coarse_from_fine(oc, of)
    ef2 = _rel_err(gc.get_f(), oc.logical_f())
    print(f"  coarse after the fill {ef2:.3e}", flush=True)
    assert ef2 < RTOL
    gc.close()
    gf.close()
