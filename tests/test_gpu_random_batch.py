"""Seeded random sets of small grids advanced as one batch (psl_batch_step)
against the same grids stepped one by one.

Each set holds 3-9 grids drawn like the random single-grid configurations
(shapes of 2-20 x 1-12 x 1-12 cells, random box-face kinds and given values,
bounceback cells, uniform or per-cell Smagorinsky constants; without periodic
copies, which a batch refuses), all in one storage.  The batch runs 7 steps as
one graph of 4 steps and 3 single steps, at each cells-per-thread and tile
order; every grid ends on the bits and shifter of its single-grid run.
"""
import numpy as np
import pytest

from tests import random_cases as rc

pytestmark = pytest.mark.gpu

SETS = 4
STEPS = 7


def _set(seed):
    rng = np.random.default_rng(7000 + seed)
    cfgs = []
    for k in range(int(rng.integers(3, 10))):
        c = rc.config(100 * seed + k)
        c["copies"] = (False, False)
        cfgs.append(c)
    return cfgs


def _start(P, c, storage):
    g = rc.grid(P, c, storage)
    P.fill_equilibrium(g, c["rho"].astype(np.float32), *c["u"].astype(np.float32))
    return g


@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_random_batches_equal_single_grid_steps(psl, storage):
    P = psl
    for seed in range(SETS):
        cfgs = _set(seed)
        singles = []
        for c in cfgs:
            g = _start(P, c, storage)
            P.updateGrid([g], STEPS)
            singles.append((g.get_f(), g.shifter.tolist()))
            g.close()
        for cpt in (1, 2):
            for order in (0, 1):
                grids = [_start(P, c, storage) for c in cfgs]
                b = P.Batch(grids)
                b.set_layout(cpt, order)
                b.steps_per_graph = 4
                b.step(STEPS)
                b.close()
                for k, (g, (f, sh)) in enumerate(zip(grids, singles)):
                    nbad = int(np.count_nonzero(g.get_f().view(np.uint32) != f.view(np.uint32)))
                    assert g.shifter.tolist() == sh, f"set {seed} grid {k}"
                    assert nbad == 0, f"set {seed} grid {k} {rc.describe(cfgs[k])}: {nbad} values differ"
                    g.close()
        print(f"[random batch] {storage} set {seed}: {len(cfgs)} grids "
              f"({sum(int(np.prod(c['shape'])) for c in cfgs)} cells) equal to single-grid steps", flush=True)
