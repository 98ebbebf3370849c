"""Boundary-condition pins: the oracle against golden vectors written by the
reference's own compiled headers (tests/golden/make_bc_golden.py), exact
moment identities, and the library's generated device code (host build)."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import oracle as O

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden", "bc_vectors.npz")

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1])
CZ = np.array([0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1])
NORMALS = [(a, b, c) for a in (-1, 0, 1) for b in (-1, 0, 1) for c in (-1, 0, 1) if (a, b, c) != (0, 0, 0)]
MOMENTS = ["000", "100", "010", "001", "200", "020", "002", "011", "101", "110", "111", "210", "201", "120", "021",
           "102", "012", "022", "202", "220", "211", "121", "112"]


def test_golden_vectors_bit_exact():
    d = np.load(GOLDEN)
    for i in range(len(d["which"])):
        code = O.normal_code(*d["normal"][i])
        f, ru = O.bc(int(d["which"][i]), code, d["f_in"][i], d["rho_u_in"][i])
        assert np.array_equal(f.view(np.uint32), d["f_out"][i].view(np.uint32)), i
        assert np.array_equal(ru.view(np.uint32), d["rho_u_out"][i].view(np.uint32)), i
    assert np.array_equal(O.bounceback(d["bb_in"]), d["bb_out"])
    assert len(d["which"]) == 26 * 4 * 16


def _unknown(n):
    return [q for q in range(27) if min(CX[q] * n[0], CY[q] * n[1], CZ[q] * n[2]) < 0]


def _meq(m, rho, u):
    """Moments of the product-form equilibrium with second moments 1/3 + u^2."""
    e = [int(ch) for ch in m]
    val = rho
    for a in range(3):
        if e[a] == 1:
            val *= u[a]
        elif e[a] == 2:
            val *= 1 / 3 + u[a] ** 2
    return val


def test_mbbc_imposes_selected_equilibrium_moments():
    """After applyMBBC, as many equilibrium moments as there are unknowns hold
    exactly; the known populations are untouched (paper eqs 2.1-2.3)."""
    import json
    tables = {t["code"]: t for t in json.load(open(os.path.join(REPO, "periodicshiftlbm_amd", "bcgen", "bc_tables.json")))}
    rng = np.random.default_rng(3)
    for n in NORMALS:
        code = O.normal_code(*n)
        t = tables[code]
        unk = _unknown(n)
        assert sorted(t["unknown"]) == unk
        assert len(t["mbbc_rows"]) == len(unk)
        for _ in range(5):
            rho = rng.uniform(0.9, 1.1)
            u = rng.uniform(-0.05, 0.05, 3)
            f = rng.uniform(0.0, 0.3, 27)
            g, _ = O.bc(0, code, f, np.array([rho, *u]))
            known = [q for q in range(27) if q not in unk]
            assert np.array_equal(g[known], f[known])
            for r in t["mbbc_rows"]:
                a, b, c = (int(ch) for ch in MOMENTS[r])
                mom = np.sum(g * CX ** a * CY ** b * CZ ** c)
                assert abs(mom - _meq(MOMENTS[r], rho, u)) < 1e-12, (n, MOMENTS[r])


def test_restore_rho_face_identity():
    """Face normals: sum_{c.n=0} f + 2 sum_{c.n=1} f = rho (1 + u.n)."""
    rng = np.random.default_rng(4)
    for n in [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]:
        cn = CX * n[0] + CY * n[1] + CZ * n[2]
        for _ in range(5):
            f = rng.uniform(0.0, 0.3, 27)
            u = rng.uniform(-0.05, 0.05, 3)
            _, ru = O.bc(3, O.normal_code(*n), f, np.array([0.0, *u]))
            lhs = f[cn == 0].sum() + 2 * f[cn == 1].sum()
            assert abs(lhs - ru[0] * (1 + np.dot(u, n))) < 1e-13


def test_restores_exact_at_rest():
    """At the rest equilibrium every restore returns rho and u = 0 exactly."""
    W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
    for n in NORMALS:
        code = O.normal_code(*n)
        f = W * 1.07
        _, ru = O.bc(1, code, f, np.zeros(4))
        assert np.allclose(ru, [1.07, 0, 0, 0], atol=1e-14)
        _, ru = O.bc(2, code, f, np.array([1.07, 9, 9, 9]))
        assert np.allclose(ru, [1.07, 0, 0, 0], atol=1e-14)
        _, ru = O.bc(3, code, f, np.array([5.0, 0, 0, 0]))
        assert abs(ru[0] - 1.07) < 1e-14
        g, _ = O.bc(0, code, f, np.array([1.07, 0, 0, 0]))
        assert np.allclose(g, f, atol=1e-15)


@pytest.fixture(scope="module")
def devmath():
    out = os.path.join(HERE, "csrc", "_build", "devmath_host.so")
    src = os.path.join(HERE, "csrc", "devmath_host.cpp")
    if not os.path.exists(out) or os.path.getmtime(out) < max(
            os.path.getmtime(src), os.path.getmtime(os.path.join(REPO, "periodicshiftlbm_amd", "csrc", "psl_math.h")),
            os.path.getmtime(os.path.join(REPO, "periodicshiftlbm_amd", "csrc", "bc_generated.inc"))):
        os.makedirs(os.path.dirname(out), exist_ok=True)
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off", "-o", out, src], check=True)
    L = ctypes.CDLL(out)
    L.dm_collide.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_float, ctypes.c_float, ctypes.c_int]
    L.dm_rescale.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_int]
    L.dm_bc.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    return L


def test_generated_device_bc_matches_golden(devmath):
    """The straight-line device BC code (host build) against the reference vectors."""
    d = np.load(GOLDEN)
    worst = 0.0
    for i in range(len(d["which"])):
        code = O.normal_code(*d["normal"][i])
        f = d["f_in"][i].copy()
        ru = d["rho_u_in"][i].copy()
        devmath.dm_bc(int(d["which"][i]), code, f.ctypes.data, ru.ctypes.data)
        ref_f, ref_ru = d["f_out"][i], d["rho_u_out"][i]
        worst = max(worst, float(np.max(np.abs(f - ref_f) / np.maximum(np.abs(ref_f), 1e-3))),
                    float(np.max(np.abs(ru - ref_ru) / np.maximum(np.abs(ref_ru), 1e-3))))
    assert worst < 1e-5, worst


def test_device_collision_formulation_matches_oracle(devmath):
    """The library's restructured FP32 collision / rescale (host build) against
    the oracle's literal restatement, FP32 and FP64."""
    rng = np.random.default_rng(0)
    n = 500
    ru = np.stack([rng.uniform(0.95, 1.05, n), *rng.uniform(-0.08, 0.08, (3, n))], 1)
    f = O.feq(ru) * (1 + rng.normal(0, 0.01, (n, 27)))
    f32 = f.astype(np.float32)
    for cs in (0.0, 0.17):
        for w_off in (0, 1):
            a = f32.copy()
            devmath.dm_collide(a.ctypes.data, n, 0.005, cs, w_off)
            o64 = O.collide(f, 0.005, cs)
            assert np.max(np.abs(a - o64) / np.abs(o64)) < 5e-6
    for c2f in (0, 1):
        a = f32.copy()
        devmath.dm_rescale(a.ctypes.data, n, c2f)
        o = O.rescale(f, bool(c2f))
        assert np.max(np.abs(a - o) / np.abs(o)) < 5e-6
