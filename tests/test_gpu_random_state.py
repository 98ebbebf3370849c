"""Seeded random restarts (psl_grid_save_state / psl_grid_load_state; no
reference counterpart): a random small configuration of tests/random_cases.py
in FP32, FP16 or BF16, stepped a random number of times, its state saved,
stepped on; a fresh grid of the same settings loads the state and takes the
same steps.  Populations and shifter must equal the uninterrupted run bit for
bit, whatever the shift reached (random step counts put every population's
shift at a different offset) and whatever the variant.
"""
import numpy as np
import pytest

import oracle as O
from tests import random_cases as rc

pytestmark = pytest.mark.gpu
CASES = 12


@pytest.mark.parametrize("seed", range(CASES))
def test_random_restart_is_bit_exact(psl, seed):
    P = psl
    c = rc.config(500 + seed)
    rng = np.random.default_rng(15000 + seed)
    storage = ["f32", "f16", "bf16"][seed % 3]
    first, then = int(rng.integers(1, 40)), int(rng.integers(1, 20))
    f0 = rc.oracle_grid(O, c, np.float32).logical_f().astype(np.float32)
    g = rc.grid(P, c, storage)
    g.tuning = int(rng.integers(0, P.TUNE_VARIANTS))
    g.set_f(f0)
    P.updateGrid([g], first)
    s = g.state()
    P.updateGrid([g], then)
    fresh = rc.grid(P, c, storage)
    fresh.tuning = int(rng.integers(0, P.TUNE_VARIANTS))
    fresh.load_state(s)
    P.updateGrid([fresh], then)
    nbad = int(np.count_nonzero(g.get_f().view(np.uint32) != fresh.get_f().view(np.uint32)))
    print(f"[random restart] seed {seed} {storage}: {rc.describe(c)}, saved after {first} steps, {then} more: "
          f"{nbad} differing values", flush=True)
    assert fresh.shifter.tolist() == g.shifter.tolist()
    assert nbad == 0
    g.close()
    fresh.close()
