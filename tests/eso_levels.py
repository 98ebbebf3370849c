"""Test helpers: esotwist populations through the write maps, and the oracle's
DIAD level recursion (updateGrid over DIADGridStruct levels,
flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:200-209)."""
from __future__ import annotations

import numpy as np

import oracle as O

OPP = np.array([0] + [q + 1 if q % 2 else q - 1 for q in range(1, 27)])


def _slots(nbr7):
    n = nbr7.shape[1]
    return np.vstack([np.arange(n, dtype=np.int64)[None], nbr7.astype(np.int64)])


def eso_get_all(raw, nbr7, flip):
    """f[q, cell] through getEsotwistWriteIndex of parity `flip`."""
    _, W = O.esotwist_tables()
    s = _slots(nbr7)
    out = np.empty_like(raw)
    for q in range(27):
        out[q] = raw[OPP[q] if flip else q, s[W[q]]]
    return out


def eso_set_all(raw, nbr7, flip, f):
    _, W = O.esotwist_tables()
    s = _slots(nbr7)
    for q in range(27):
        raw[OPP[q] if flip else q, s[W[q]]] = f[q]


class OracleLevel:
    def __init__(self, d, flags, bcvals, nu, cs, dtype=np.float32):
        self.n = d["n"]
        self.ijk = np.ascontiguousarray(d["ijk"], np.int32)
        self.nbr7 = np.ascontiguousarray(d["nbr"][:7], np.int32)
        self.info = O.make_info(d["nx"], d["ny"], d["nz"], nu=nu, cs=cs)
        self.flags = np.ascontiguousarray(flags, np.uint8)
        self.bcvals = np.ascontiguousarray(bcvals, dtype).reshape(-1)
        self.f = np.zeros((27, self.n), dtype)
        self.flip = False
        self.tables = {k: np.ascontiguousarray(d[k], np.int32) for k in ("c2f_read", "c2f_write", "f2c_write",
                                                                       "f2c_read")}
        self.moving = None  # (cells, u3): moving-wall cells after each update
        self.forced = None  # (cells, ut3): forced-velocity cells after each update

    def get(self):
        return eso_get_all(self.f, self.nbr7, self.flip)

    def set(self, f):
        eso_set_all(self.f, self.nbr7, self.flip, np.asarray(f, self.f.dtype))


def communicate(c: OracleLevel, f: OracleLevel):
    t = c.tables
    O.eso_to_coarse(c.f, c.nbr7, c.flip, f.f, f.nbr7, f.flip, t["f2c_write"], t["f2c_read"])
    O.eso_to_fine(c.f, c.nbr7, c.flip, f.f, f.nbr7, f.flip, t["c2f_write"], t["c2f_read"])


def update_grid(levels, l=0):
    L = levels[l]
    L.flip = not L.flip
    O.esotwist_update(L.f, L.nbr7, L.ijk, L.info, L.flags, L.bcvals, L.flip)
    if L.moving is not None:
        O.eso_moving_bounceback(L.f, L.nbr7, L.flip, *L.moving)
    if L.forced is not None:
        R = L.f.dtype.type
        O.eso_forced_velocity(L.f, L.nbr7, L.flip, *L.forced, R(L.info.nu), R(L.info.cs))
    if l < len(levels) - 1:
        update_grid(levels, l + 1)
        update_grid(levels, l + 1)
        communicate(L, levels[l + 1])
