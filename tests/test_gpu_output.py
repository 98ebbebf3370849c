"""Section cuts on the device against the fields of the grid's own populations.

A single periodic-shift grid (a sphere channel of 64 x 32 x 32 after 20
steps) is cut in all three planes and compared with the same plane assembled
on the host from psl_get_f: FP64 moments, the conversion p = (rho - 1)
rho_nominal c^2 and u res / 1000 / dtPhys, the plane's velocity components and
the bounceback marker (1e-5 relative on p's rho - 1 part through an absolute
bound, 1e-5 on velocities).  Wall-refined DIAD levels are cut likewise and
compared with a host painter that applies the same rules (levels coarse to
fine, coarse cells over their footprint, the largest cell index where finer
cells share a pixel).
"""
import numpy as np
import pytest

import oracle as O
from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RHO0, CS = 1.225, 3.4


def _moments(f):
    c = np.array(O.velocity_set(), np.float64)
    f = f.astype(np.float64)
    rho = f.sum(0)
    return rho, (c @ f) / rho


def _components(plane, ux, uy, uz):
    return {"XY": (ux, uy, uz), "ZY": (uz, uy, ux), "ZX": (uz, ux, uy)}[plane]


def test_section_cut_grid(psl):
    P = psl
    shape = (64, 32, 32)
    nx, ny, nz = shape
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = ((i3 - 16.0) ** 2 + (j3 - 15.5) ** 2 + (k3 - 15.5) ** 2 <= 16.0).reshape(-1)
    flags = cases.channel_flags(shape, solid)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=0.01, res=2.0, dtPhys=0.5)
    g = P.GridStruct(info, "f32")
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(0.04))
    g.set_periodic_copies(y=True)
    P.fill_equilibrium(g)
    P.updateGrid([g], 20)
    f = g.get_f()
    rho, u = _moments(f)
    vel = (info.res / 1000.0) / info.dtPhys
    for plane, cut in (("XY", 16), ("ZY", 20), ("ZX", 15)):
        got = P.section_cut(g, plane, cut, RHO0, CS)
        cell = (k3 * ny + j3) * nx + i3
        sel = {"XY": cell[cut, :, :], "ZY": cell[:, :, cut].T, "ZX": cell[:, cut, :].T}[plane]
        p = (rho[sel] - 1.0) * RHO0 * CS * CS
        uh, uv, un = _components(plane, *(u[a][sel] * vel for a in range(3)))
        marker = ((flags[sel] & 7) == cases.BOUNCEBACK).astype(np.float32)
        print(f"[cut] {plane} {cut}: {got.shape}", flush=True)
        assert got.shape == sel.shape + (5,)
        assert np.max(np.abs(got[..., 0] - p)) < 1e-5 * RHO0 * CS * CS
        for a, want in ((1, uh), (2, uv), (3, un)):
            assert np.max(np.abs(got[..., a] - want)) < 1e-5 * vel
        assert np.array_equal(got[..., 4], marker)
    # getAvgInletPressure: the mean density over the j = 0 plane as a pressure
    p_in = P.getAvgInletPressure(g, RHO0, CS)
    assert abs(p_in - (rho[cell[:, 0, :]].mean() - 1.0) * RHO0 * CS * CS) < 1e-5 * RHO0 * CS * CS
    # export3DPlot: every cell, (i, j, k) order with k fastest
    vol = P.export3DPlot(g, 3, path=None, rho_nominal=RHO0, soundspeed=CS)
    assert vol.shape == (nx, ny, nz, 5)
    order = cell.transpose(2, 1, 0)  # [i, j, k] -> cell
    assert np.max(np.abs(vol[..., 0] - (rho[order] - 1.0) * RHO0 * CS * CS)) < 1e-5 * RHO0 * CS * CS
    for a in range(3):
        assert np.max(np.abs(vol[..., 1 + a] - u[a][order] * vel)) < 1e-5 * vel
    assert np.array_equal(vol[..., 4], ((flags[order] & 7) == cases.BOUNCEBACK).astype(np.float32))
    g.close()


def _paint(levels, plane, cut, target_scale, nlevels, h, v):
    out = np.zeros((v, h, 6), np.float64)
    out[..., 4] = 1.0
    for l, (g, lv_info, ijk, flags) in enumerate(levels):
        scale = 1 << (nlevels - l - 1)
        i, j, k = (a.astype(np.int64) * scale for a in ijk)
        ic, ih, iv = {"XY": (k, i, j), "ZY": (i, k, j), "ZX": (j, k, i)}[plane]
        on = (cut >= ic) & (cut < ic + scale)
        rho, u = _moments(g.get_f())
        vel = (lv_info.res / 1000.0) / lv_info.dtPhys
        span = max(1, scale // target_scale)
        winner = np.full((v, h), -1, np.int64)
        for cell in np.flatnonzero(on):
            y0, x0 = iv[cell] // target_scale, ih[cell] // target_scale
            blk = winner[y0:y0 + span, x0:x0 + span]
            np.maximum(blk, cell, out=blk)
        ys, xs = np.nonzero(winner >= 0)
        cells = winner[ys, xs]
        uh, uv, un = _components(plane, *(u[a][cells] * vel for a in range(3)))
        out[ys, xs, 0] = (rho[cells] - 1.0) * RHO0 * CS * CS
        out[ys, xs, 1], out[ys, xs, 2], out[ys, xs, 3] = uh, uv, un
        out[ys, xs, 4] = (flags[cells] & 7) == cases.BOUNCEBACK
        out[ys, xs, 5] = l
    return out


@pytest.mark.parametrize("plane,cut", [("XY", 40), ("ZY", 33), ("ZX", 47)])
def test_section_cut_diad_levels(psl, plane, cut):
    P = psl
    lv, grids = cases.diad_sphere(P, "f32", n0=16, levels=3, span=1, radius=2.0)
    for g in grids:
        g.fill_equilibrium(ux=0.04)
    P.updateGrid(grids, 3)
    got = P.section_cut(grids, plane, cut, RHO0, CS)
    levels = []
    for l, g in enumerate(grids):
        li, _ = lv.info(l)
        flags = cases.diad_sphere_flags(lv.array(l, "ijk"), lv.array(l, "bounceback"),
                                        (li.cellCountX, li.cellCountY, li.cellCountZ), l)
        levels.append((g, li, lv.array(l, "ijk"), flags))
    v, h = got.shape[:2]
    want = _paint(levels, plane, cut, 1, 3, h, v)
    print(f"[cut] DIAD {plane} {cut}: {got.shape}, levels on the plane "
          f"{np.unique(got[..., 5]).astype(int).tolist()}", flush=True)
    assert np.array_equal(got[..., 5], want[..., 5].astype(np.float32))
    assert np.array_equal(got[..., 4], want[..., 4].astype(np.float32))
    assert np.max(np.abs(got[..., 0] - want[..., 0])) < 1e-5 * RHO0 * CS * CS
    assert np.max(np.abs(got[..., 1:4] - want[..., 1:4])) < 1e-5 * 1.0
    # a pixel budget below the finest plane: the middle level sets the pixel size
    budget = got.shape[0] * got.shape[1]
    small = P.section_cut(grids, plane, cut, RHO0, CS, max_pixels=budget)
    assert small.shape[0] * 2 == got.shape[0] and small.shape[1] * 2 == got.shape[1]
    want2 = _paint(levels, plane, cut, 2, 3, small.shape[1], small.shape[0])
    assert np.array_equal(small[..., 5], want2[..., 5].astype(np.float32))
    assert np.array_equal(small[..., 4], want2[..., 4].astype(np.float32))
    assert np.max(np.abs(small[..., 1:4] - want2[..., 1:4])) < 1e-5 * 1.0


def test_flow_report_diad_levels(psl):
    """getFlowReportXY over DIAD levels: the plane clipped to physical bounds,
    then means over the pixels outside walls, against the host painter."""
    P = psl
    lv, grids = cases.diad_sphere(P, "f32", n0=16, levels=3, span=1, radius=2.0)
    for g in grids:
        g.fill_equilibrium(ux=0.04)
    P.updateGrid(grids, 3)
    bounds = np.array([-6.0, -4.0, -100.0, 8.0, 5.0, 100.0], np.float32)
    cut = 32
    rep = P.getFlowReport(grids, "XY", cut, bounds)
    fi, _ = lv.info(2)
    f32 = np.float32
    i0 = max(0, int(np.floor((bounds[0] - f32(fi.ox)) / f32(fi.res))))
    i1 = min(fi.cellCountX, int(np.ceil((bounds[3] - f32(fi.ox)) / f32(fi.res))))
    j0 = max(0, int(np.floor((bounds[1] - f32(fi.oy)) / f32(fi.res))))
    j1 = min(fi.cellCountY, int(np.ceil((bounds[4] - f32(fi.oy)) / f32(fi.res))))
    levels = []
    for l, g in enumerate(grids):
        li, _ = lv.info(l)
        flags = cases.diad_sphere_flags(lv.array(l, "ijk"), lv.array(l, "bounceback"),
                                        (li.cellCountX, li.cellCountY, li.cellCountZ), l)
        # lattice units: a level info with res / 1000 / dtPhys = 1
        raw = P.InfoStruct(res=1000.0, dtPhys=1.0)
        levels.append((g, raw, lv.array(l, "ijk"), flags))
    full = _paint(levels, "XY", cut, 1, 3, fi.cellCountX, fi.cellCountY)
    win = full[j0:j1, i0:i1]
    open_px = win[..., 4] == 0
    rho = win[..., 0][open_px] / (RHO0 * CS * CS) + 1.0
    # _paint stores the plane components (XY: ux, uy, uz) already
    want = {"ux": win[..., 1][open_px].mean(), "uy": win[..., 2][open_px].mean(),
            "uz": win[..., 3][open_px].mean(), "rho": (rho - 1.0).mean() + 1.0,
            "areamm2": open_px.sum() * float(fi.res) ** 2}
    print(f"[flow report] {rep}\n  host {want}", flush=True)
    assert open_px.sum() > 0 and (~open_px).sum() > 0
    for key in ("ux", "uy", "uz", "rho"):
        assert abs(rep[key] - want[key]) < 1e-6, key
    assert rep["areamm2"] == pytest.approx(want["areamm2"], rel=1e-12)


def test_section_cut_scalar_transport(psl, tmp_path):
    """The scalar-transport cut: the five components of the plain cut and, as a
    sixth, sum f T / sum f of each cell, against the host from the grid's own
    f and T (1e-5 relative); written as six components for the scalar plotter."""
    P = psl
    shape = nx, ny, nz = (48, 24, 24)
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = ((i3 - 12.0) ** 2 + (j3 - 11.5) ** 2 + (k3 - 11.5) ** 2 <= 16.0).reshape(-1)
    flags = cases.channel_flags(shape, solid)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=0.02, res=2.0, dtPhys=0.5)
    g = P.GridStruct(info, "f32")
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(0.04))
    P.fill_equilibrium(g, ux=np.full(g.n, 0.04, np.float32))
    st = P.ScalarTransportStruct(g, 0.7)
    given = np.full(g.n, np.nan, np.float32)
    given[i3.reshape(-1) == 0] = 1.0
    st.set_given(given)
    for _ in range(15):
        P.applyStreaming(g)
        P.applyScalarTransportLocalCellUpdate(g, st)
    f = g.get_f().astype(np.float64)
    T = st.get_T().astype(np.float64)
    conc = (f * T).sum(0) / f.sum(0)
    assert conc.max() > 0.5
    cell = (k3 * ny + j3) * nx + i3
    for plane, cut in (("XY", 11), ("ZY", 3), ("ZX", 12)):
        got = P.section_cut(g, plane, cut, RHO0, CS, scalar_transport=True)
        plain = P.section_cut(g, plane, cut, RHO0, CS)
        sel = {"XY": cell[cut, :, :], "ZY": cell[:, :, cut].T, "ZX": cell[:, cut, :].T}[plane]
        print(f"[cut] scalar {plane} {cut}: {got.shape}, max scalar {got[..., 5].max():.3f}", flush=True)
        assert got.shape == sel.shape + (6,)
        assert np.array_equal(got[..., :5], plain)
        assert np.max(np.abs(got[..., 5] - conc[sel])) < 1e-5 * max(1.0, np.abs(conc).max())
    path = str(tmp_path / "sim_data.bin")
    data = P.exportSectionCutPlot(g, "ZY", 3, 7, path=path, rho_nominal=RHO0, soundspeed=CS, scalar_transport=True)
    with open(path, "rb") as fh:
        hdr = np.fromfile(fh, dtype=np.int32, count=4)
        body = np.fromfile(fh, dtype=np.float32).reshape((hdr[1], hdr[2], hdr[3]))
    assert list(hdr) == [7, ny, nz, 6] and np.array_equal(body, data)
    g.close()


def test_toilet_paper_cut(psl):
    """The unrolled cylinder about z: row v at angle v res / r reads the cell
    getIJKCellIndexFromXYZ names, column k; p, uz, the tangential and radial
    velocity and the marker against host moments of the grid's populations.
    A radius whose circle leaves the grid is rejected."""
    P = psl
    nx, ny, nz = 40, 36, 12
    res = 2.0
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x3, y3 = i3 * res - 39.0, j3 * res - 35.0
    solid = ((x3 - 8.0) ** 2 + y3 ** 2 <= 30.0).reshape(-1)
    flags = cases.channel_flags((nx, ny, nz), solid)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=0.02, res=res, dtPhys=0.5,
                        ox=-39.0, oy=-35.0, oz=0.0)
    g = P.GridStruct(info, "f32")
    g.set_flags(flags)
    g.set_given_values(cases.channel_given_values(0.04))
    P.fill_equilibrium(g, ux=np.full(g.n, 0.04, np.float32))
    P.updateGrid([g], 10)
    rho, u = _moments(g.get_f())
    vel = (res / 1000.0) / info.dtPhys
    for r in (12.5, 21.0):
        got = P.exportSectionCutPlotToiletPaperZ(g, r, 1, path=None, rho_nominal=RHO0, soundspeed=CS)
        rows = int(np.float32(2.0 * np.float32(3.14159265359) * np.float32(r)) / np.float32(res))
        assert got.shape == (rows, nz, 5)
        v = np.arange(rows, dtype=np.float32)
        th = (v * np.float32(res)) / np.float32(r)
        x, y = np.float32(r) * np.cos(th), np.float32(r) * np.sin(th)
        i = ((x - np.float32(-39.0)) / np.float32(res) + np.float32(0.5)).astype(np.int64)
        j = ((y - np.float32(-35.0)) / np.float32(res) + np.float32(0.5)).astype(np.int64)
        cell = np.arange(nz)[None, :] * nx * ny + (j * nx + i)[:, None]
        c, s = np.cos(th)[:, None], np.sin(th)[:, None]
        ux, uy, uz = (u[a][cell] * vel for a in range(3))
        print(f"[cut] toilet paper r = {r}: {got.shape}, marker pixels {int(got[..., 4].sum())}", flush=True)
        assert np.max(np.abs(got[..., 0] - (rho[cell] - 1.0) * RHO0 * CS * CS)) < 1e-5 * RHO0 * CS * CS
        assert np.max(np.abs(got[..., 1] - uz)) < 1e-5 * vel
        assert np.max(np.abs(got[..., 2] - (-ux * s + uy * c))) < 2e-5 * vel
        assert np.max(np.abs(got[..., 3] - (ux * c + uy * s))) < 2e-5 * vel
        assert np.array_equal(got[..., 4], ((flags[cell] & 7) == cases.BOUNCEBACK).astype(np.float32))
    with pytest.raises(P.PslError):
        P.exportSectionCutPlotToiletPaperZ(g, 60.0, 1, path=None)
    g.close()


def _c_div(a, b):
    return np.fix(a / b).astype(np.int64)  # C integer division, towards zero


def _paint_cyl(levels, r, target_scale, nlevels, h, v, fres, io, jo):
    """Host painter of the unrolled cylinder over DIAD levels with the device's
    rules, FP32 geometry as the reference computes it."""
    f32 = np.float32
    out = np.zeros((v, h, 6), np.float64)
    out[..., 4] = 1.0
    for l, (g, lv_info, ijk, flags) in enumerate(levels):
        scale = 1 << (nlevels - l - 1)
        i, j, k = (a.astype(np.int64) * scale for a in ijk)
        xc = (i.astype(f32) + f32(0.5 * scale) - f32(io)) * f32(fres)
        yc = (j.astype(f32) + f32(0.5 * scale) - f32(jo)) * f32(fres)
        dist = np.sqrt(xc * xc + yc * yc)
        half = f32(scale * fres * 0.5)
        near = ~((dist < f32(r) - half * f32(1.5)) | (dist > f32(r) + half * f32(1.5)))
        th = np.arctan2(yc, xc)
        th = np.where(th < 0, th + f32(2.0) * f32(3.14159265359), th).astype(f32)
        fv = (th * f32(r) / f32(fres)).astype(np.int64)
        span = max(1, scale // target_scale)
        y0 = _c_div(fv - (span * target_scale) // 2, target_scale)
        x0 = k // target_scale
        rho, u = _moments(g.get_f())
        vel = (lv_info.res / 1000.0) / lv_info.dtPhys
        winner = np.full((v, h), -1, np.int64)
        for cell in np.flatnonzero(near):
            for dy in range(span):
                y = (y0[cell] + dy) % v
                xs = x0[cell] + np.arange(span)
                xs = xs[xs < h]
                winner[y, xs] = np.maximum(winner[y, xs], cell)
        ys, xs = np.nonzero(winner >= 0)
        cells = winner[ys, xs]
        t = (ys.astype(f32) * f32(target_scale) * f32(fres)) / f32(r)
        c, s = np.cos(t).astype(np.float64), np.sin(t).astype(np.float64)
        ux, uy, uz = (u[a][cells] * vel for a in range(3))
        out[ys, xs, 0] = (rho[cells] - 1.0) * RHO0 * CS * CS
        out[ys, xs, 1], out[ys, xs, 2], out[ys, xs, 3] = uz, -ux * s + uy * c, ux * c + uy * s
        out[ys, xs, 4] = (flags[cells] & 7) == cases.BOUNCEBACK
        out[ys, xs, 5] = l
    return out


def test_toilet_paper_cut_diad_levels(psl):
    """The unrolled cylinder over three DIAD levels against the host painter:
    levels and markers on all but a few pixels (cells whose angle or distance
    falls on a rounding boundary may land one pixel apart between the device's
    and numpy's atan2), fields within 1e-5 where the painters agree."""
    P = psl
    lv, grids = cases.diad_sphere(P, "f32", n0=16, levels=3, span=1, radius=2.0)
    for g in grids:
        g.fill_equilibrium(ux=0.04)
    P.updateGrid(grids, 3)
    levels = []
    for l, g in enumerate(grids):
        li, _ = lv.info(l)
        flags = cases.diad_sphere_flags(lv.array(l, "ijk"), lv.array(l, "bounceback"),
                                        (li.cellCountX, li.cellCountY, li.cellCountZ), l)
        levels.append((g, li, lv.array(l, "ijk"), flags))
    fin, _ = lv.info(2)
    io = int(np.float32(0.0 - fin.ox) / np.float32(fin.res) + np.float32(0.5))
    jo = int(np.float32(0.0 - fin.oy) / np.float32(fin.res) + np.float32(0.5))
    for r in (2.6, 4.1):
        got = P.exportSectionCutPlotToiletPaperZ(grids, r, 1, path=None, rho_nominal=RHO0, soundspeed=CS)
        v, h = got.shape[:2]
        assert v == int(np.float32(2.0 * np.float32(3.14159265359) * np.float32(r)) / np.float32(fin.res))
        assert h == fin.cellCountZ
        want = _paint_cyl(levels, r, 1, 3, h, v, fin.res, io, jo)
        same = (got[..., 5] == want[..., 5]) & (got[..., 4] == want[..., 4]) & \
            (np.abs(got[..., 0] - want[..., 0]) < 1e-5 * RHO0 * CS * CS)
        print(f"[cut] DIAD toilet paper r = {r}: {got.shape}, levels {np.unique(got[..., 5]).astype(int).tolist()}, "
              f"{int((~same).sum())} of {same.size} pixels differ", flush=True)
        assert len(np.unique(got[..., 5])) >= 2
        assert (~same).sum() <= 0.005 * same.size
        assert np.max(np.abs(got[same][:, 1:4] - want[same][:, 1:4])) < 1e-5
    # a pixel budget the finest plane exceeds: the middle level sets the pixel size
    r = 4.1
    got = P.exportSectionCutPlotToiletPaperZ(grids, r, 1, path=None, rho_nominal=RHO0, soundspeed=CS, max_pixels=2000)
    mid, _ = lv.info(1)
    v, h = got.shape[:2]
    assert v == int(np.float32(2.0 * np.float32(3.14159265359) * np.float32(r)) / np.float32(mid.res))
    assert h == mid.cellCountZ
    want = _paint_cyl(levels, r, 2, 3, h, v, fin.res, io, jo)
    same = (got[..., 5] == want[..., 5]) & (got[..., 4] == want[..., 4])
    print(f"[cut] DIAD toilet paper r = {r} at the middle level's pixels: {got.shape}, "
          f"{int((~same).sum())} of {same.size} pixels differ", flush=True)
    assert (~same).sum() <= 0.005 * same.size
    assert np.max(np.abs(got[same][:, 1:4] - want[same][:, 1:4])) < 1e-5


def test_cut_markers_of_moving_and_forced_cells(psl):
    """Moving-wall and forced cells are flagged bounceback here; the reference's
    plot marker is its bounceback marker alone, so both paint as open in the
    cuts, while its flow reports count forced cells as blocked
    (include/flowReportFunctions.h:163).  Against the host painter with the
    flags the reference would see."""
    P = psl
    lv, grids = cases.diad_sphere(P, "f32", n0=16, levels=3, span=1, radius=2.0)
    for g in grids:
        g.fill_equilibrium(ux=0.04)
    P.updateGrid(grids, 3)
    cut = 32
    rng = np.random.default_rng(6)
    levels_plot, levels_report = [], []
    for l, g in enumerate(grids):
        li, _ = lv.info(l)
        ijk = lv.array(l, "ijk")
        flags = cases.diad_sphere_flags(ijk, lv.array(l, "bounceback"), (li.cellCountX, li.cellCountY, li.cellCountZ), l)
        plot_flags, report_flags = flags.copy(), flags.copy()
        if l == 2:
            on = np.flatnonzero((ijk[2] == cut) & (flags == cases.FLUID))
            pick = rng.choice(on, 60, replace=False)
            moving, forced = np.sort(pick[:30]).astype(np.int32), np.sort(pick[30:]).astype(np.int32)
            dev = flags.copy()
            dev[pick] = cases.BOUNCEBACK
            g.set_flags(dev)
            g.set_moving_bounceback(moving, np.zeros((30, 3), np.float32))
            g.set_forced_velocity(forced, np.zeros((30, 3), np.float32))
            report_flags[forced] = cases.BOUNCEBACK
        levels_plot.append((g, li, ijk, plot_flags))
        levels_report.append((g, P.InfoStruct(res=1000.0, dtPhys=1.0), ijk, report_flags))
    got = P.section_cut(grids, "XY", cut, RHO0, CS)
    v, h = got.shape[:2]
    want = _paint(levels_plot, "XY", cut, 1, 3, h, v)
    assert np.array_equal(got[..., 4], want[..., 4].astype(np.float32))
    assert np.array_equal(got[..., 5], want[..., 5].astype(np.float32))
    bounds = np.array([-6.0, -4.0, -100.0, 8.0, 5.0, 100.0], np.float32)
    rep = P.getFlowReport(grids, "XY", cut, bounds)
    fi, _ = lv.info(2)
    f32 = np.float32
    i0 = max(0, int(np.floor((bounds[0] - f32(fi.ox)) / f32(fi.res))))
    i1 = min(fi.cellCountX, int(np.ceil((bounds[3] - f32(fi.ox)) / f32(fi.res))))
    j0 = max(0, int(np.floor((bounds[1] - f32(fi.oy)) / f32(fi.res))))
    j1 = min(fi.cellCountY, int(np.ceil((bounds[4] - f32(fi.oy)) / f32(fi.res))))
    win = _paint(levels_report, "XY", cut, 1, 3, fi.cellCountX, fi.cellCountY)[j0:j1, i0:i1]
    open_px = win[..., 4] == 0
    print(f"[cut] moving/forced markers: plot open pixels {int((got[..., 4] == 0).sum())}, "
          f"report open pixels {int(open_px.sum())}", flush=True)
    assert rep["areamm2"] == pytest.approx(open_px.sum() * float(fi.res) ** 2, rel=1e-12)
    assert abs(rep["ux"] - win[..., 1][open_px].mean()) < 1e-6
