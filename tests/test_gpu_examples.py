"""The reference's nested-box sphere (flowAroundSphere/flowAroundSphereMainSingle.cu):
four periodic-shift levels through the level recursion, against the oracle's
recursion, and the C++ example driver written against include/psl_facade.hpp
against the same case driven from Python.

* Four levels (the driver's gridLevelCount) at 44 coarse cells per edge
  (levels of 44^3, 26^3, 32^3 and 38^3 cells; ghost box faces on the finer
  levels, the sphere inside all four), FP32: every level's f within 1e-5
  relative of the FP32 oracle after 6 coarse steps (48 finest steps), and the
  same at 36 cells (levels of 36^3, 20^3, 26^3 and 32^3).
* The example binary (built in-tree by build()) at 44 cells for 200
  iterations: its level sizes equal the Python layout and its drag-coefficient
  history (historyData.bin) equals the Python path's within 1e-5 relative:
  both call the same library on the same flags, so the forces are the same.
"""
import os
import re
import struct
import subprocess
import tempfile

import numpy as np
import pytest

from periodicshiftlbm_amd import build as B
from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

GIVEN = np.array([[1.0, 0.07, 0.0, 0.0]] + [[1.0, 0.0, 0.0, 0.0]] * 3, np.float32)
CS = 0.1


def _info(P, d):
    return P.InfoStruct(**{k: (float(v) if isinstance(v, np.floating) else v) for k, v in d.items()})


def _gpu_levels(P, levels):
    grids = []
    for d, flags in levels:
        g = P.GridStruct(_info(P, d), "f32")
        g.set_flags(flags)
        g.set_given_values(GIVEN)
        g.set_smagorinsky(CS)
        grids.append(g)
    return grids


def _sub(d):
    if "iSubgridStart" not in d:
        return None
    return (d["iSubgridStart"], d["jSubgridStart"], d["kSubgridStart"], d["iSubgridEnd"], d["jSubgridEnd"],
            d["kSubgridEnd"])


@pytest.mark.parametrize("nlev,cells", [(4, 44), (4, 36)])
def test_nested_sphere_levels_match_oracle(psl, oracle_mod, nlev, cells):
    P, O = psl, oracle_mod
    levels, _ = cases.reference_sphere_levels(nlev, cells)
    print(f"[nested sphere] {nlev} levels, {cells} coarse cells: "
          f"{[(d['cellCountX'], d['cellCountY'], d['cellCountZ']) for d, _ in levels]}", flush=True)
    grids = _gpu_levels(P, levels)
    ors = []
    for (d, flags), g in zip(levels, grids):
        og = O.Grid(d["cellCountX"], d["cellCountY"], d["cellCountZ"], dtype=np.float32, nu=float(d["nu"]), cs=CS,
                    flags=flags, bcvals=GIVEN, sub=_sub(d))
        n = flags.size
        # flow along x everywhere except inside the sphere, so every level moves from the first step
        ux = np.where(flags == cases.BOUNCEBACK, 0.0, 0.07)
        og.fill_equilibrium(np.ones(n), ux, np.zeros(n), np.zeros(n))
        g.set_f(og.logical_f().astype(np.float32))
        ors.append(og)

    def oracle_update(level):
        ors[level].step()
        if level < nlev - 1:
            for _ in range(2):
                oracle_update(level + 1)
            O.coarse_fine_comm(ors[level], ors[level + 1])

    steps = 6
    P.updateGrid(grids, steps)
    for _ in range(steps):
        oracle_update(0)
    for lev, (g, og) in enumerate(zip(grids, ors)):
        a = g.get_f().astype(np.float64)
        b = og.logical_f().astype(np.float64)
        err = float(np.max(np.abs(a - b) / np.maximum(np.abs(b), 1e-3)))
        print(f"  level {lev}: max rel err {err:.3e}", flush=True)
        assert err < 1e-5
    for g in grids:
        g.close()


def _read_history(path):
    data = open(path, "rb").read()
    (count,) = struct.unpack("<i", data[:4])
    return np.frombuffer(data[4:4 + 4 * count], np.float32)


def test_sphere_example_driver_matches_python_path(psl):
    P = psl
    exe = os.path.join(B.EXAMPLE_BIN, "flow_around_sphere")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is not built (run __graft_entry__.build())")
    cells, iters, chunk = 44, 200, 100
    print(f"[sphere example] {exe} {cells} {iters} {chunk}", flush=True)
    with tempfile.TemporaryDirectory() as d:
        r = subprocess.run([exe, str(cells), str(iters), str(chunk), d], capture_output=True, text=True, timeout=300)
        print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
        assert r.returncode == 0
        hist = _read_history(os.path.join(d, "historyData.bin"))
        assert os.path.getsize(os.path.join(d, "sim_data.bin")) > 0
    counts = [int(x) for x in re.findall(r"Cell count on grid \d+: (\d+)", r.stdout)]
    levels, (lo, hi) = cases.reference_sphere_levels(4, cells)
    assert counts == [fl.size for _, fl in levels]
    assert hist.size == iters + 1
    # the same case from Python
    grids = _gpu_levels(P, levels)
    for g in grids:
        P.fill_equilibrium(g)
    cd = np.zeros(iters + 1)
    infos = [d for d, _ in levels]
    for it in range(iters + 1):
        P.updateGrid(grids, 1)
        fx = float(grids[-1].momentum_exchange(lo, hi)[0])
        cd[it] = cases.reference_sphere_drag_coefficient(fx, infos)
        if it % chunk == 0 and it != 0:
            for lev in range(len(grids) - 2, -1, -1):
                P.fillCoarseGridFromFine(grids[lev], grids[lev + 1])
    print(f"  C_D at {iters}: example {hist[-1]:.6f} python {cd[-1]:.6f}", flush=True)
    scale = np.max(np.abs(cd[10:]))
    assert scale > 0
    assert np.max(np.abs(hist.astype(np.float64) - cd)) <= 1e-5 * scale
    for g in grids:
        g.close()
