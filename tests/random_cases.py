"""Seeded random small periodic-shift configurations shared by the GPU tests
(test_gpu_random_configs, test_gpu_random_batch).

A case is a shape of 2-20 x 1-12 x 1-12 cells, a kind and a given-value slot
per box face, random given values, 5-15 % bounceback cells inside, a uniform or
per-cell Smagorinsky constant or none, periodic copies on the axes without
listed faces, and a random near-equilibrium state.  Listed face cells drop the
normal component of every axis whose faces are not listed; axes under three
cells thick have no listed faces.
"""
import numpy as np

from periodicshiftlbm_amd import cases

FACE_KINDS = [cases.GIVEN_U, cases.GIVEN_RHO, cases.GIVEN_NONE, cases.GIVEN_RHOU, cases.BOUNCEBACK, cases.FLUID]


def config(seed, large=False):
    """Small grids (2-20 x 1-12 x 1-12 cells), or with large=True 24-72 cells
    per axis: many workgroups, wrap thresholds inside waves, several blocks of
    staged boundary records."""
    rng = np.random.default_rng((31000 if large else 1000) + seed)
    if large:
        shape = tuple(int(x) for x in rng.integers(24, 73, 3))
    else:
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


def grid(P, c, storage):
    nx, ny, nz = c["shape"]
    g = P.GridStruct(P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=c["nu"]), storage)
    g.set_flags(c["flags"])
    g.set_given_values(c["bc"])
    g.set_smagorinsky(c["cs"] if c["cs_cell"] is None else c["cs_cell"])
    g.set_periodic_copies(x=c["copies"][0], y=c["copies"][1])
    return g


def oracle_grid(O, c, dtype):
    nx, ny, nz = c["shape"]
    og = O.Grid(nx, ny, nz, dtype=dtype, nu=c["nu"], cs=c["cs"], flags=c["flags"], bcvals=c["bc"],
                cs_cell=c["cs_cell"])
    og.fill_equilibrium(c["rho"], *c["u"])
    og.f[:] = og.logical_f() * (1 + c["noise"])
    return og


def describe(c):
    kinds = np.bincount(c["flags"] & 7, minlength=8)
    return (f"shape={c['shape']} fluid={kinds[0]} bounceback={kinds[1]} listed={int(kinds[2:6].sum())} "
            f"cs={'per-cell' if c['cs_cell'] is not None else c['cs']:} copies={c['copies']}")
