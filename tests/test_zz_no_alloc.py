"""The no-allocation momentum exchange (psl_reduce_momentum_exchange_noalloc)
against the allocating entry point it mirrors, through the C ABI.

Both run the same kernels with the same block count, so their FP64 results
must be identical bit for bit; the no-allocation entry point is also held to
the oracle at the sphere-channel tolerance.  Its scratch is one fixed array
per device that keeps the previous call's partials, so two calls in a row are
checked, the second on a smaller box (fewer blocks than the first wrote)."""
import ctypes

import numpy as np
import pytest

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu


def _sphere_channel(P, O, shape, storage):
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = ((i - 12) ** 2 + (j - 12) ** 2 + (k - 12) ** 2) <= 25
    flags = cases.channel_flags(shape, solid.reshape(-1))
    bc = cases.channel_given_values(0.04)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=0.0048)
    g = P.GridStruct(info, storage)
    g.set_flags(flags)
    g.set_given_values(bc)
    g.set_smagorinsky(0.17)
    og = O.Grid(nx, ny, nz, dtype=np.float32, nu=0.0048, cs=0.17, flags=flags, bcvals=bc)
    n = nx * ny * nz
    og.fill_equilibrium(np.ones(n), np.zeros(n), np.zeros(n), np.zeros(n))
    g.set_f(og.logical_f().astype(np.float32))
    g.set_periodic_copies(y=True)
    return g, og


def _call(P, fn, g, lo, hi):
    out = np.full(3, np.nan, np.float64)
    lo_a = None if lo is None else np.ascontiguousarray(lo, np.int32)
    hi_a = None if hi is None else np.ascontiguousarray(hi, np.int32)
    rc = fn(g._h, None if lo_a is None else lo_a.ctypes.data, None if hi_a is None else hi_a.ctypes.data,
            out.ctypes.data)
    return rc, out


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_noalloc_momentum_exchange_equals_allocating_entry_point(psl, oracle_mod, storage):
    P, O = psl, oracle_mod
    L = P.lib()
    shape = (48, 24, 24)
    print(f"[no-alloc momentum exchange] storage={storage} shape={shape}", flush=True)
    g, og = _sphere_channel(P, O, shape, storage)
    steps = 30
    P.updateGrid([g], steps)
    for _ in range(steps):
        og.step(periodic_y=True)
    # first call: the whole grid (108 blocks); second: a box around the sphere (9 blocks)
    for lo, hi in ((None, None), ([6, 6, 6], [19, 19, 19])):
        rc_a, ref = _call(P, L.psl_reduce_momentum_exchange, g, lo, hi)
        rc_b, got = _call(P, L.psl_reduce_momentum_exchange_noalloc, g, lo, hi)
        assert rc_a == 0 and rc_b == 0, L.psl_last_error()
        print(f"  box {lo}..{hi}: allocating {ref}, no-alloc {got}")
        assert np.array_equal(got, ref)
        if storage == "f32":
            fo = og.momentum_exchange(lo or [0, 0, 0], hi or list(shape))
            assert np.all(np.abs(got - fo) <= 1e-5 * np.max(np.abs(fo)))
    # the Python mirror's flag reaches the same entry point
    assert np.array_equal(g.momentum_exchange([6, 6, 6], [19, 19, 19], noalloc=True), ref)
    # an empty box returns zeros without a launch
    rc, out = _call(P, L.psl_reduce_momentum_exchange_noalloc, g, [5, 5, 5], [5, 9, 9])
    assert rc == 0 and np.array_equal(out, np.zeros(3))
    g.close()


def test_noalloc_momentum_exchange_rejects_boxes_beyond_its_scratch(psl):
    """A box of more than PSL_NOALLOC_MAX_BOX_CELLS cells is refused on the
    host (no kernel runs); the allocating entry point takes it."""
    P = psl
    L = P.lib()
    limit = 16777216
    nx, ny, nz = 257, 256, 256  # 16,842,752 cells
    assert nx * ny * nz > limit
    print(f"[no-alloc limit] shape={(nx, ny, nz)}", flush=True)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=0.01)
    g = P.GridStruct(info, "f16")
    flags = np.zeros(nx * ny * nz, np.uint8)
    flags[: nx * ny] = cases.BOUNCEBACK
    g.set_flags(flags)
    P.fill_equilibrium(g)
    rc, out = _call(P, L.psl_reduce_momentum_exchange_noalloc, g, None, None)
    assert rc != 0
    assert b"exceeds the scratch limit" in L.psl_last_error()
    # just inside the limit along z: accepted, and equal to the allocating call
    hi = [nx, ny, limit // (nx * ny)]
    rc_a, ref = _call(P, L.psl_reduce_momentum_exchange, g, None, hi)
    rc_b, got = _call(P, L.psl_reduce_momentum_exchange_noalloc, g, None, hi)
    assert rc_a == 0 and rc_b == 0
    assert np.array_equal(got, ref)
    g.close()
