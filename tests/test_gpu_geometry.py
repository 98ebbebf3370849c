"""GPU parity of block refinement transfers, the level recursion and the STL
voxelizer against the CPU oracle.

Refinement (FP32): distributions within 1e-5 relative after the transfers and
after several coarse steps of the two-level recursion.
Voxelizer: masks bit-exact; edge counts exact.
"""
import os
import tempfile

import numpy as np
import pytest

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b), 1e-3)))


def _two_level(P, O, nu=0.01, cs=0.0):
    """Coarse 24 x 20 x 18 periodic box (copies in x, y) with a nested fine box."""
    nxc, nyc, nzc = 24, 20, 18
    sub = (6, 5, 4, 18, 15, 13)
    nxf, nyf, nzf = 2 * (sub[3] - sub[0]), 2 * (sub[4] - sub[1]), 2 * (sub[5] - sub[2])
    ci = P.InfoStruct(cellCountX=nxc, cellCountY=nyc, cellCountZ=nzc, nu=nu, res=1.0,
                      iSubgridStart=sub[0], jSubgridStart=sub[1], kSubgridStart=sub[2],
                      iSubgridEnd=sub[3], jSubgridEnd=sub[4], kSubgridEnd=sub[5])
    nu_f = 2 * nu  # same physical viscosity on the finer grid (dt and dx halved)
    fi = P.InfoStruct(gridID=1, cellCountX=nxf, cellCountY=nyf, cellCountZ=nzf, nu=nu_f, res=0.5,
                      ox=sub[0] - 0.25, oy=sub[1] - 0.25, oz=sub[2] - 0.25)
    gc = P.GridStruct(ci, "f32")
    gf = P.GridStruct(fi, "f32")
    k, j, i = np.meshgrid(np.arange(nzf), np.arange(nyf), np.arange(nxf), indexing="ij")
    ghost = (i == 0) | (i == nxf - 1) | (j == 0) | (j == nyf - 1) | (k == 0) | (k == nzf - 1)
    fflags = np.where(ghost, cases.GHOST, cases.FLUID).astype(np.uint8).reshape(-1)
    gf.set_flags(fflags)
    gc.set_periodic_copies(x=True, y=True)
    gc.set_smagorinsky(cs)
    gf.set_smagorinsky(cs)
    oc = O.Grid(nxc, nyc, nzc, dtype=np.float32, nu=nu, cs=cs, sub=sub)
    of = O.Grid(nxf, nyf, nzf, dtype=np.float32, nu=nu_f, cs=cs, flags=fflags)

    def field(nx, ny, nz, x0, y0, z0, h):
        k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        x = 2 * np.pi * (x0 + i * h) / 23
        y = 2 * np.pi * (y0 + j * h) / 19
        z = 2 * np.pi * (z0 + k * h) / 18
        ux = 0.03 * np.sin(x) * np.cos(y) * np.cos(z)
        uy = -0.03 * np.cos(x) * np.sin(y) * np.cos(z)
        return np.ones(nx * ny * nz), ux.reshape(-1), uy.reshape(-1), np.zeros(nx * ny * nz)

    oc.fill_equilibrium(*field(nxc, nyc, nzc, 0, 0, 0, 1.0))
    of.fill_equilibrium(*field(nxf, nyf, nzf, sub[0] - 0.25, sub[1] - 0.25, sub[2] - 0.25, 0.5))
    gc.set_f(oc.logical_f())
    gf.set_f(of.logical_f())
    return gc, gf, oc, of


@pytest.mark.parametrize("sub", [(6, 5, 4, 7, 15, 13), (6, 5, 4, 18, 6, 13), (6, 5, 4, 18, 15, 5)])
def test_sub_grid_one_coarse_cell_wide_is_refused(psl, sub):
    """A sub-grid one coarse cell wide along an axis the coarse grid extends in
    would put its fine-to-coarse ring outside the sub-grid (fine cells past the
    fine box); the transfers, the coarse fill and the recursion refuse it
    before launching a transfer."""
    P = psl
    ci = P.InfoStruct(cellCountX=24, cellCountY=20, cellCountZ=18, nu=0.01, res=1.0,
                      iSubgridStart=sub[0], jSubgridStart=sub[1], kSubgridStart=sub[2],
                      iSubgridEnd=sub[3], jSubgridEnd=sub[4], kSubgridEnd=sub[5])
    fi = P.InfoStruct(gridID=1, cellCountX=2 * (sub[3] - sub[0]), cellCountY=2 * (sub[4] - sub[1]),
                      cellCountZ=2 * (sub[5] - sub[2]), nu=0.02, res=0.5)
    gc, gf = P.GridStruct(ci, "f32"), P.GridStruct(fi, "f32")
    print(f"[thin sub-grid] {sub}", flush=True)
    for call in (lambda: P.applyCoarseFineGridCommunication(gc, gf), lambda: P.fillCoarseGridFromFine(gc, gf),
                 lambda: P.updateGrid([gc, gf], 1)):
        with pytest.raises(RuntimeError, match="two coarse cells"):
            call()
    gc.close()
    gf.close()


def test_coarse_fine_communication_matches_oracle(psl, oracle_mod):
    P, O = psl, oracle_mod
    gc, gf, oc, of = _two_level(P, O)
    print("[refine] coarse-fine communication", flush=True)
    P.applyCoarseFineGridCommunication(gc, gf)
    O.coarse_fine_comm(oc, of)
    ec, ef = _rel_err(gc.get_f(), oc.logical_f()), _rel_err(gf.get_f(), of.logical_f())
    print(f"  coarse {ec:.3e} fine {ef:.3e}")
    assert ec < RTOL and ef < RTOL
    P.fillCoarseGridFromFine(gc, gf)
    O.fill_coarse_from_fine(oc, of)
    assert _rel_err(gc.get_f(), oc.logical_f()) < RTOL


def test_two_level_recursion_matches_oracle(psl, oracle_mod):
    """updateGrid recursion (flowAroundSphereMainSingle.cu:158-167) over two levels."""
    P, O = psl, oracle_mod
    gc, gf, oc, of = _two_level(P, O, nu=0.01, cs=0.12)
    steps = 4
    print(f"[refine] two-level recursion, {steps} coarse steps", flush=True)
    P.updateGrid([gc, gf], steps)
    for _ in range(steps):
        oc.step(periodic_x=True, periodic_y=True)
        for _ in range(2):
            of.step()
        O.coarse_fine_comm(oc, of)
    ec, ef = _rel_err(gc.get_f(), oc.logical_f()), _rel_err(gf.get_f(), of.logical_f())
    print(f"  coarse {ec:.3e} fine {ef:.3e}")
    assert ec < RTOL and ef < RTOL
    # covered coarse cells were skipped, as in the reference
    bulk, nbc, skipped = gc.census()
    assert skipped == (12 - 4) * (10 - 4) * (9 - 4)


def _physical(logical, shift):
    n = logical.shape[1]
    out = np.empty_like(logical)
    idx = np.arange(n, dtype=np.int64)
    for q in range(27):
        out[q, (idx + int(shift[q])) % n] = logical[q]
    return out


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_coarse_fine_communication_wide_faces(psl, oracle_mod, storage):
    """Transfers on a sub-grid whose faces span several workgroups (80 x 30 x 30
    coarse cells), touching the coarse grid's y = 0 edge (no fine face there,
    one-sided differences in the interpolation), with wrapped population
    shifts on both levels.  FP32 within 1e-5 relative; FP16 storage against
    the oracle in FP64 on the stored (rounded) inputs, within 1e-4 absolute
    (FP16 resolves f - w to 2^-11 relative, |f - w| < 0.05 here)."""
    P, O = psl, oracle_mod
    nxc, nyc, nzc = 100, 40, 36
    sub = (2, 0, 3, 82, 30, 33)
    nxf, nyf, nzf = 2 * (sub[3] - sub[0]), 2 * (sub[4] - sub[1]), 2 * (sub[5] - sub[2])
    ci = P.InfoStruct(cellCountX=nxc, cellCountY=nyc, cellCountZ=nzc, nu=0.01, res=1.0,
                      iSubgridStart=sub[0], jSubgridStart=sub[1], kSubgridStart=sub[2],
                      iSubgridEnd=sub[3], jSubgridEnd=sub[4], kSubgridEnd=sub[5])
    fi = P.InfoStruct(gridID=1, cellCountX=nxf, cellCountY=nyf, cellCountZ=nzf, nu=0.02, res=0.5,
                      ox=sub[0] - 0.25, oy=sub[1] - 0.25, oz=sub[2] - 0.25)
    gc = P.GridStruct(ci, storage)
    gf = P.GridStruct(fi, storage)
    k, j, i = np.meshgrid(np.arange(nzf), np.arange(nyf), np.arange(nxf), indexing="ij")
    ghost = (i == 0) | (i == nxf - 1) | (j == 0) | (j == nyf - 1) | (k == 0) | (k == nzf - 1)
    fflags = np.where(ghost, cases.GHOST, cases.FLUID).astype(np.uint8).reshape(-1)
    gf.set_flags(fflags)
    dt = np.float32 if storage == "f32" else np.float64
    oc = O.Grid(nxc, nyc, nzc, dtype=dt, nu=0.01, sub=sub)
    of = O.Grid(nxf, nyf, nzf, dtype=dt, nu=0.02, flags=fflags)
    rng = np.random.default_rng(11)
    for g, o, (nx, ny, nz), (x0, y0, z0, h) in ((gc, oc, (nxc, nyc, nzc), (0, 0, 0, 1.0)),
                                                (gf, of, (nxf, nyf, nzf), (sub[0] - 0.25, sub[1] - 0.25,
                                                                           sub[2] - 0.25, 0.5))):
        kk, jj, ii = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        x = 2 * np.pi * (x0 + ii * h) / 37
        y = 2 * np.pi * (y0 + jj * h) / 29
        z = 2 * np.pi * (z0 + kk * h) / 31
        rho = 1.0 + 0.01 * np.sin(x + 2 * y) * np.cos(z)
        ux = 0.04 * np.sin(x) * np.cos(y) * np.cos(z)
        uy = -0.03 * np.cos(x) * np.sin(y) * np.cos(2 * z)
        uz = 0.02 * np.sin(x + y + z)
        tmp = O.Grid(nx, ny, nz, dtype=np.float64)
        tmp.fill_equilibrium(rho.reshape(-1), ux.reshape(-1), uy.reshape(-1), uz.reshape(-1))
        shift = rng.integers(0, nx * ny * nz, 27).astype(np.int32)
        g.shifter = shift
        g.set_f(tmp.f.astype(np.float32))
        stored = g.get_f()  # the inputs as the storage holds them
        o.shift[:] = shift
        o.f[:] = _physical(stored.astype(dt), shift)
    print(f"[refine] wide faces, {storage} storage", flush=True)
    P.applyCoarseFineGridCommunication(gc, gf)
    O.coarse_fine_comm(oc, of)
    if storage == "f32":
        ec, ef = _rel_err(gc.get_f(), oc.logical_f()), _rel_err(gf.get_f(), of.logical_f())
        print(f"  coarse {ec:.3e} fine {ef:.3e}")
        assert ec < RTOL and ef < RTOL
    else:
        ec = float(np.max(np.abs(gc.get_f() - oc.logical_f())))
        ef = float(np.max(np.abs(gf.get_f() - of.logical_f())))
        print(f"  coarse {ec:.3e} fine {ef:.3e} (absolute)")
        assert ec < 1e-4 and ef < 1e-4
    # the fine faces were rewritten except the skipped y = 0 face (FP32 resolves
    # the second-order interpolation error of this smooth field)
    changed = np.any(gf.get_f() != stored, axis=0).reshape(nzf, nyf, nxf)
    assert not changed[2:-2, 0, 2:-2].any()
    if storage == "f32":
        assert changed[:, 1:, 0].all() and changed[:, 1:, -1].all() and changed[0, 1:, :].all()
        assert changed[:, -1, :].all()
    gc.close()
    gf.close()


def test_voxelizer_matches_oracle_bit_exact(psl, oracle_mod):
    P, O = psl, oracle_mod
    tri = cases.icosphere(3, 7.3, (15.2, 13.7, 12.9))
    info = P.InfoStruct(cellCountX=32, cellCountY=28, cellCountZ=26, res=1.0)
    stl = P.STLStruct.from_triangles(tri)
    m = P.BoolArray(32 * 28 * 26)
    print("[stl] icosphere subdivision 3", flush=True)
    P.applyMarkersInsideSTL(m, stl, True, info)
    got = m.get()
    want = O.markers_inside_stl(tri, (0, 0, 0), (0, 0, 0), 1.0, 32, 28, 26)
    assert np.array_equal(got, want)
    assert got.sum() > 1000
    # inverted inside value, grid origin and resolution other than 1
    info2 = P.InfoStruct(cellCountX=40, cellCountY=36, cellCountZ=30, res=0.75, ox=1.5, oy=-0.25, oz=0.5)
    stl.set_origin(0.5, 0.25, -0.5)
    P.applyMarkersInsideSTL(m2 := P.BoolArray(40 * 36 * 30), stl, False, info2)
    want2 = O.markers_inside_stl(tri, (0.5, 0.25, -0.5), (1.5, -0.25, 0.5), 0.75, 40, 36, 30, inside_value=0)
    assert np.array_equal(m2.get(), want2)


def test_sphere_config_voxelization(psl, oracle_mod):
    """The sphere configuration: icosphere subdivision 4 (5120 triangles), D = 32."""
    P, O = psl, oracle_mod
    tri = cases.icosphere(4, cases.SPHERE_DIAMETER / 2, cases.SPHERE_CENTER)
    assert tri.shape == (5120, 9)
    stl = P.STLStruct.from_triangles(tri)
    bad, counts = P.checkSTLEdges(stl)
    obad, ocounts = O.check_stl_edges(tri)
    assert bad == 0 and obad == 0 and np.array_equal(counts, ocounts)
    shape = (128, 128, 128)
    info = P.InfoStruct(cellCountX=shape[0], cellCountY=shape[1], cellCountZ=shape[2], res=1.0)
    m = P.BoolArray(int(np.prod(shape)))
    P.applyMarkersInsideSTL(m, stl, True, info)
    got = m.get()
    want = O.markers_inside_stl(tri, (0, 0, 0), (0, 0, 0), 1.0, *shape)
    assert np.array_equal(got, want)
    vol = got.sum()
    print(f"  solid cells {vol}, sphere volume {4 / 3 * np.pi * 16 ** 3:.0f}")
    assert abs(vol / (4 / 3 * np.pi * 16 ** 3) - 1) < 0.02


def test_stl_file_round_trip_and_mask_ops(psl, oracle_mod):
    P, O = psl, oracle_mod
    tri = cases.icosphere(2, 5.0, (8.0, 8.0, 8.0))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "ico.stl")
        cases.write_binary_stl(path, tri)
        stl = P.readSTL(path)
        assert stl.triangleCount == tri.shape[0]
        assert np.array_equal(stl.triangles(), tri)
        b = stl.bounds
        assert np.allclose(b[:3], tri.reshape(-1, 3).min(0)) and np.allclose(b[3:], tri.reshape(-1, 3).max(0))
    with pytest.raises(P.PslError):
        P.readSTL("/nonexistent/file.stl")
    n = 16 ** 3
    a, bm, out = P.BoolArray(n), P.BoolArray(n), P.BoolArray(n)
    rng = np.random.default_rng(0)
    ha, hb = rng.integers(0, 2, n).astype(np.uint8), rng.integers(0, 2, n).astype(np.uint8)
    a.set(ha)
    bm.set(hb)
    P.multiplyBoolArrays(a, bm, out)
    assert np.array_equal(out.get(), ha & hb)
    P.sumBoolArrays(a, bm, out)
    assert np.array_equal(out.get(), ha | hb)
    P.invertBoolArray(out)
    assert np.array_equal(out.get(), 1 - (ha | hb))
    # open mesh reported
    stl_open = P.STLStruct.from_triangles(tri[1:])
    bad, _ = P.checkSTLEdges(stl_open)
    assert bad == 3


def test_bounceback_mask_applies_to_flags(psl, oracle_mod):
    P, O = psl, oracle_mod
    shape = (20, 18, 16)
    info = P.InfoStruct(cellCountX=20, cellCountY=18, cellCountZ=16, nu=0.01)
    g = P.GridStruct(info)
    flags = cases.channel_flags(shape)
    g.set_flags(flags)
    tri = cases.icosphere(2, 4.0, (10.0, 9.0, 8.0))
    m = P.BoolArray(int(np.prod(shape)))
    P.applyMarkersInsideSTL(m, P.STLStruct.from_triangles(tri), True, info)
    P.applyBouncebackMask(g, m)
    got = g.get_flags()
    want = flags.copy()
    want[m.get().astype(bool)] = cases.BOUNCEBACK
    assert np.array_equal(got, want)
