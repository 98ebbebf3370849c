"""Seeded random periodic boxes through the esotwist (DIAD) update.

Each case is a fully periodic box of 1-12 cells per axis (one- and two-cell
axes included, where a cell is its own +i neighbour), 5-15 % bounceback cells,
a random Smagorinsky constant and a random near-equilibrium state, 7 steps (so
the last one ends on the odd flipper):

* FP32 in the builder's (k, j, i) cell order against the oracle's restatement
  of the reference maps, within 1e-5 on the fluid cells (the tolerance of the
  other esotwist tests);
* FP32, FP16 and BF16: the same box with its cells relabelled by a random
  permutation (neighbour columns and flags relabelled with it) gives every cell
  the same bits.  A per-cell update cannot depend on where a cell is stored;
  under a random order almost no +i, +ij, +ik, +ijk neighbour is the next cell,
  so the update reads its neighbour columns where in (k, j, i) order it mostly
  takes the adjacent cell.
"""
import numpy as np
import pytest

from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu

RTOL = 1e-5
CASES = 12
STEPS = 7


def _case(seed):
    rng = np.random.default_rng(5000 + seed)
    shape = tuple(int(x) for x in rng.integers(1, 13, 3))
    n = int(np.prod(shape))
    flags = np.zeros(n, np.uint8)
    flags[rng.random(n) < rng.uniform(0.05, 0.15)] = cases.BOUNCEBACK
    return dict(shape=shape, n=n, flags=flags, cs=float(rng.choice([0.0, rng.uniform(0.05, 0.2)])),
                nu=float(rng.uniform(0.005, 0.05)), rho=rng.uniform(0.99, 1.01, n).astype(np.float32),
                u=rng.uniform(-0.02, 0.02, (3, n)).astype(np.float32), perm=rng.permutation(n))


def _run(P, c, storage, relabel, f0=None):
    nx, ny, nz = c["shape"]
    ijk, nbr = P.DIADGridStruct.dense_box(nx, ny, nz)
    flags, rho, u = c["flags"], c["rho"], c["u"]
    if relabel:
        p = c["perm"]  # cell c is stored at p[c]
        ijk2, nbr2 = np.empty_like(ijk), np.empty_like(nbr)
        ijk2[:, p] = ijk
        nbr2[:, p] = p[nbr]
        ijk, nbr = ijk2, nbr2
        f2 = np.empty_like(flags)
        f2[p] = flags
        flags = f2
        r2, u2 = np.empty_like(rho), np.empty_like(u)
        r2[p] = rho
        u2[:, p] = u
        rho, u = r2, u2
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=c["nu"])
    g = P.DIADGridStruct(info, ijk, nbr, storage)
    g.set_flags(flags)
    g.set_smagorinsky(c["cs"])
    if f0 is None:
        g.fill_equilibrium(rho, *u)
    else:
        g.set_f(f0)
    for _ in range(STEPS):
        P.applyStreaming(g)
        P.applyLocalCellUpdate(g)
    f = g.get_f()
    flip = g.esotwistFlipper
    g.close()
    if relabel:
        f = f[:, c["perm"]]
    return f, flip


@pytest.mark.parametrize("seed", range(CASES))
def test_random_box_matches_oracle(psl, oracle_mod, seed):
    P, O = psl, oracle_mod
    c = _case(seed)
    nx, ny, nz = c["shape"]
    n = c["n"]
    print(f"[random esotwist] seed {seed}: shape={c['shape']} bounceback={int((c['flags'] == 1).sum())} "
          f"cs={c['cs']:.3f}", flush=True)
    ijk, nbr = P.DIADGridStruct.dense_box(nx, ny, nz)
    rd, wr = O.esotwist_tables()
    # the oracle array after fill_equilibrium (flipper false), as the update reads it
    feq = O.feq(np.stack([c["rho"], *c["u"]], 1).astype(np.float32))
    A = np.zeros((27, n), np.float32)
    slots = np.vstack([np.arange(n)[None], nbr])
    for q in range(27):
        A[q, slots[wr[q]]] = feq[:, q]
    opp = np.array([0] + [q + 1 if q % 2 else q - 1 for q in range(1, 27)])

    def logical(flipper):
        out = np.empty_like(A)
        for q in range(27):
            out[q] = A[opp[q] if flipper else q, slots[wr[q]]]
        return out

    # both start from the same bits
    got, flip = _run(P, c, "f32", False, f0=logical(False))
    assert flip is True  # seven flips from false
    oinfo = O.make_info(nx, ny, nz, c["nu"], c["cs"])
    bc = np.zeros((4, 4), np.float32)
    flipper = False
    for _ in range(STEPS):
        flipper = not flipper
        O.esotwist_update(A, nbr, ijk, oinfo, c["flags"], bc, flipper)
    want = logical(flipper)
    fluid = c["flags"] == 0
    err = float(np.max(np.abs(got[:, fluid].astype(np.float64) - want[:, fluid]) /
                       np.maximum(np.abs(want[:, fluid]), 1e-3))) if fluid.any() else 0.0
    print(f"  max rel err {err:.3e}", flush=True)
    assert err < RTOL


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
def test_random_boxes_independent_of_cell_order(psl, storage):
    P = psl
    for seed in range(CASES):
        c = _case(seed)
        a, fa = _run(P, c, storage, False)
        b, fb = _run(P, c, storage, True)
        nbad = int(np.count_nonzero(a.view(np.uint32) != b.view(np.uint32)))
        print(f"[esotwist order] {storage} seed {seed} shape={c['shape']}: {nbad} differing values", flush=True)
        assert fa == fb
        assert np.all(np.isfinite(a))
        assert nbad == 0
