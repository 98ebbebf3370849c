"""Dense-array restatement of buildDIADGrids for tiny boxes (test helper).

include/DIADFunctions.h:640-1046 works on sorted cell lists with binary-search
neighbour lookup, sort-based compaction, child generation and index maps.
Here every level is a dense boolean box (nz, ny, nx) with an `exists` mask:
neighbour marking is a padded 26-shift OR, compaction is C-order nonzero,
children are a 2x upsampling, index maps are dense lookups and the periodic
esotwist neighbour is a vectorised walk.  None of the list machinery of the
oracle is shared, so the two agree only if both orders, searches and maps are
right.  Marker rules and the STL inside test are evaluated over the dense box
(the latter by the oracle's whole-grid voxelizer, pinned on its own).
"""
from __future__ import annotations

import numpy as np

DIRS26 = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1) if (dx, dy, dz) != (0, 0, 0)]
ESO_DIRS = [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, 0, 1), (0, 1, 1), (1, 1, 1), (-1, 0, 0), (0, -1, 0),
            (0, 0, -1)]
f32 = np.float32


def _dilate(src, exists):
    """out[c] = some existing 26-neighbour of c is set in src (non-periodic)."""
    nz, ny, nx = src.shape
    p = np.pad(src & exists, 1)
    out = np.zeros_like(src)
    for dx, dy, dz in DIRS26:
        out |= p[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]
    return out


def _contains(rule, info, K, J, I):
    shape = int(rule["shape"])
    p = rule["p"].astype(f32)
    if shape == 4:
        def rng(v, lo, hi, n):
            lo = lo + n if lo < 0 else lo
            hi = hi + n if hi < 0 else hi
            return (v >= lo) & (v <= hi)
        ix = rule["idx"]
        return rng(I, ix[0], ix[1], info["nx"]) & rng(J, ix[2], ix[3], info["ny"]) & rng(K, ix[4], ix[5], info["nz"])
    x = I.astype(f32) * f32(info["res"]) + f32(info["ox"])
    y = J.astype(f32) * f32(info["res"]) + f32(info["oy"])
    z = K.astype(f32) * f32(info["res"]) + f32(info["oz"])
    if shape == 3:
        return (x > p[0]) & (x < p[3]) & (y > p[1]) & (y < p[4]) & (z > p[2]) & (z < p[5])
    dx, dy, dz = x - p[0], y - p[1], z - p[2]
    w2 = p[3] * p[3]
    if shape == 2:
        return (dx > f32(0)) & (dx < p[4]) & ((dy * dy + dz * dz) < w2)
    r2 = dx * dx + dy * dy + dz * dz
    return r2 <= w2 if shape == 0 else r2 < w2


def _rules(rules, what, info, m):
    nz, ny, nx = m.shape
    K, J, I = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    m = m.copy()
    for r in rules:
        if int(r["what"]) == what and (int(r["level_mask"]) >> info["gridID"]) & 1:
            m[_contains(r, info, K, J, I)] = bool(r["value"])
    return m


def _stl_union(oracle, stls, info, shape):
    m = np.zeros(shape, bool)
    for tri, origin in stls:
        s = oracle.markers_inside_stl(tri, origin, (info["ox"], info["oy"], info["oz"]), info["res"], info["nx"],
                                      info["ny"], info["nz"])
        m |= s.reshape(shape).astype(bool)
    return m


def _index_map(exists):
    idx = np.full(exists.shape, -1, np.int64)
    flat = np.flatnonzero(exists.ravel())
    idx.ravel()[flat] = np.arange(flat.size)
    return idx


def _eso_nbrs(exists):
    nz, ny, nx = exists.shape
    idx = _index_map(exists)
    K, J, I = [a.ravel()[np.flatnonzero(exists.ravel())] for a in
               np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")]
    out = np.full((10, K.size), -1, np.int64)
    for d, (dx, dy, dz) in enumerate(ESO_DIRS):
        todo = np.ones(K.size, bool)
        s = 0
        while todo.any():
            s += 1
            cand = idx[(K + s * dz) % nz, (J + s * dy) % ny, (I + s * dx) % nx]
            hit = todo & (cand >= 0)
            out[d, hit] = cand[hit]
            todo &= ~hit
    return out


def build(oracle, base, levels, span, nu_phys, stls=(), rules=None):
    rules = [] if rules is None else list(rules)
    info = dict(base, gridID=0)
    exists = np.ones((base["nz"], base["ny"], base["nx"]), bool)
    enfB = enfF = None
    out = []
    for level in range(levels):
        shape = exists.shape
        if level == levels - 1:
            bb = _stl_union(oracle, stls, info, shape)
            bb = _rules(rules, 0, info, bb)
            if levels > 1:
                bb |= enfB
            fluid = ~bb
            if levels > 1:
                fluid |= enfF
            keep = (fluid | _dilate(fluid, exists)) & exists
            d = dict(info, exists=keep, bounceback=~fluid, interface=np.zeros(shape, bool))
            out.append(d)
            break
        # markFinestFluidBounceback
        fluid = np.zeros(shape, bool)
        fbb = np.zeros(shape, bool)
        for tl in range(level, levels):
            ns = 1 << (tl - level)
            tres = f32(info["res"]) / f32(ns)
            start = f32(-0.5) * f32(info["res"]) + f32(0.5) * tres
            for sx in range(ns):
                for sy in range(ns):
                    for sz in range(ns):
                        si = dict(info, ox=f32(f32(info["ox"]) + start) + f32(sx) * tres,
                                  oy=f32(f32(info["oy"]) + start) + f32(sy) * tres,
                                  oz=f32(f32(info["oz"]) + start) + f32(sz) * tres)
                        m = _rules(rules, 0, si, _stl_union(oracle, stls, si, shape))
                        fbb |= m
                        fluid |= ~m
        if level:
            fluid |= enfF
        border = _dilate(fluid, exists) & ~fluid
        keep = fluid | border
        ref = fbb | border
        new = np.zeros(shape, bool)
        for _ in range(span + levels - level - 1):
            new |= _dilate(ref, exists)
            new |= ref
            ref = new & keep
        ref = _rules(rules, 1, info, ref)
        f2c = _dilate(ref, exists) & keep & ~ref
        c2f = _dilate(f2c, exists) & keep & ~f2c & ~ref
        ref = ref | f2c | c2f
        cbb = _rules(rules, 0, info, np.zeros(shape, bool)) | _stl_union(oracle, stls, info, shape)
        under = c2f | f2c
        keep = ((fluid | border) & ~ref) | c2f | f2c
        if level:
            keep |= enfF
        keep &= exists
        ref &= exists
        up = lambda a: a.repeat(2, 0).repeat(2, 1).repeat(2, 2)
        nxt_exists, nxt_enfB, nxt_enfF = up(ref), up(under & cbb & ref), up(under & ~cbb & ref)
        bb = _rules(rules, 0, info, np.zeros(shape, bool)) | _stl_union(oracle, stls, info, shape)
        if level:
            bb = (bb | enfB) & ~enfF
        outlet = _rules(rules, 2, info, np.zeros(shape, bool))
        out.append(dict(info, exists=keep, bounceback=bb, interface=(f2c | c2f), c2f=c2f & ~bb, f2c=f2c & ~bb,
                        outlet_m=outlet))
        res = f32(info["res"]) * f32(0.5)
        dt = f32(info["dt"]) * f32(0.5)
        info = dict(gridID=level + 1, res=res, dt=dt, nu=(dt * f32(nu_phys)) / ((res / f32(1000)) * (res / f32(1000))),
                    ox=f32(info["ox"]) - res * f32(0.5), oy=f32(info["oy"]) - res * f32(0.5),
                    oz=f32(info["oz"]) - res * f32(0.5), nx=2 * info["nx"], ny=2 * info["ny"], nz=2 * info["nz"])
        exists, enfB, enfF = nxt_exists, nxt_enfB, nxt_enfF
    # lists
    result = []
    for lev, d in enumerate(out):
        ex = d["exists"]
        flat = np.flatnonzero(ex.ravel())
        nz, ny, nx = ex.shape
        k, rem = np.divmod(flat, ny * nx)
        j, i = np.divmod(rem, nx)
        r = {key: d[key] for key in ("res", "ox", "oy", "oz", "nx", "ny", "nz")}
        r["n"] = flat.size
        r["ijk"] = np.stack([i, j, k]).astype(np.int32)
        r["bounceback"] = d["bounceback"].ravel()[flat].astype(np.uint8)
        r["interface"] = d["interface"].ravel()[flat].astype(np.uint8)
        r["nbr"] = _eso_nbrs(ex).astype(np.int32)
        if lev < len(out) - 1:
            fine_idx = _index_map(out[lev + 1]["exists"])
            for name, coarse_name, fine_name in (("c2f", "c2f_read", "c2f_write"), ("f2c", "f2c_write", "f2c_read")):
                sel = np.flatnonzero(d[name].ravel()[flat])
                r[coarse_name] = sel.astype(np.int32)
                r[fine_name] = fine_idx[2 * k[sel], 2 * j[sel], 2 * i[sel]].astype(np.int32)
            r["outlet"] = np.flatnonzero(d["outlet_m"].ravel()[flat]).astype(np.int32)
        else:
            for name in ("c2f_read", "c2f_write", "f2c_write", "f2c_read", "outlet"):
                r[name] = np.zeros(0, np.int32)
        result.append(r)
    return result
