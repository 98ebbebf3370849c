"""Write tests/golden/bc_vectors.npz from the reference's compiled boundary headers.

Run in the development container only (it needs the reference tree to build
oracle/_ref/libpsl_ref_bc.so via `make -C oracle ref`).  The output holds
inputs and outputs only: for every outer normal of the D3Q27 box (26 codes)
and each of applyMBBC / restoreRhoUxUyUz / restoreUxUyUz / restoreRho, a batch
of near-equilibrium FP32 populations and given (rho, u), plus a batch for
applyBounceback.  tests/test_oracle_bc.py checks the oracle against it.
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.abspath(os.path.join(HERE, "..", ".."))

W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)
CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1])
CZ = np.array([0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1])
NORMALS = [[1, 0, 0], [0, 1, 0], [0, 0, 1], [-1, 0, 0], [0, -1, 0], [0, 0, -1],
           [0, 1, 1], [1, 0, 1], [1, 1, 0], [0, -1, -1], [-1, 0, -1], [-1, -1, 0],
           [0, 1, -1], [1, 0, -1], [1, -1, 0], [0, -1, 1], [-1, 0, 1], [-1, 1, 0],
           [1, 1, 1], [-1, -1, -1], [-1, 1, 1], [1, -1, 1], [1, 1, -1], [1, -1, -1], [-1, 1, -1], [-1, -1, 1]]


def feq(rho, u):
    cu = CX * u[0] + CY * u[1] + CZ * u[2]
    u2 = (u ** 2).sum()
    return rho * W * (1 + 3 * cu + 4.5 * cu * cu - 1.5 * u2)


def main():
    lib_path = os.path.join(REPO, "oracle", "_ref", "libpsl_ref_bc.so")
    lib = ctypes.CDLL(lib_path)
    lib.ref_bc_f32.argtypes = [ctypes.c_int] * 4 + [ctypes.c_void_p, ctypes.c_void_p]
    rng = np.random.default_rng(20261202)
    per = 16
    f_in, ru_in, f_out, ru_out, which_a, normal_a = [], [], [], [], [], []
    for n in NORMALS:
        for which in range(4):
            for _ in range(per):
                rho = rng.uniform(0.9, 1.1)
                u = rng.uniform(-0.08, 0.08, size=3)
                f = (feq(rho, u) * (1 + rng.normal(0, 0.02, size=27))).astype(np.float32)
                ru = np.array([rho, *u], dtype=np.float32)
                fo, ruo = f.copy(), ru.copy()
                rc = lib.ref_bc_f32(which, n[0], n[1], n[2], fo.ctypes.data, ruo.ctypes.data)
                assert rc == 0
                f_in.append(f); ru_in.append(ru); f_out.append(fo); ru_out.append(ruo)
                which_a.append(which); normal_a.append(n)
    bb_in = rng.uniform(0.0, 0.3, size=(per, 27)).astype(np.float32)
    bb_out = bb_in.copy()
    dummy = np.zeros(4, np.float32)
    for i in range(per):
        row = bb_out[i].copy()
        lib.ref_bc_f32(4, 0, 0, 0, row.ctypes.data, dummy.ctypes.data)
        bb_out[i] = row
    out = os.path.join(HERE, "bc_vectors.npz")
    np.savez_compressed(out, f_in=np.array(f_in), rho_u_in=np.array(ru_in), f_out=np.array(f_out),
                        rho_u_out=np.array(ru_out), which=np.array(which_a, np.int32),
                        normal=np.array(normal_a, np.int32), bb_in=bb_in, bb_out=bb_out)
    print("wrote", out, len(f_in), "vectors")


if __name__ == "__main__":
    sys.exit(main())
