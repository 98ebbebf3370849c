"""ctypes bindings for the CPU oracle (oracle/libpsl_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg as the checker.  The shipped library never loads it.

Arrays follow the reference's storage: distributions are q-major SoA,
f[q * N + slot], slot = (cell + shift[q]) mod N, cell = k*Nx*Ny + j*Nx + i
(include/types.h:42-45, include/cellFunctions.h:18-21, include/applyStreaming.h:33-41).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsl_oracle.so")

# flag byte (shared with the product's C ABI)
FLUID, BOUNCEBACK, GIVEN_RHO, GIVEN_U, GIVEN_NONE, GIVEN_RHOU, GHOST, COVERED = range(8)
PERIODIC_X, PERIODIC_Y, PERIODIC_Z = 8, 16, 32


class OracleInfo(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("i_sub_start", ctypes.c_int), ("j_sub_start", ctypes.c_int), ("k_sub_start", ctypes.c_int),
                ("i_sub_end", ctypes.c_int), ("j_sub_end", ctypes.c_int), ("k_sub_end", ctypes.c_int),
                ("nu", ctypes.c_double), ("cs", ctypes.c_double)]


def make_info(nx, ny, nz, nu=0.0, cs=0.0, sub=None) -> OracleInfo:
    s = sub or (0, 0, 0, 0, 0, 0)
    return OracleInfo(nx, ny, nz, s[0], s[1], s[2], s[3], s[4], s[5], float(nu), float(cs))


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    P = ctypes.c_void_p
    L.oracle_build_flags.restype = ctypes.c_char_p
    L.oracle_build_flags.argtypes = []
    L.oracle_velocity_set.argtypes = [P, P, P]
    for suf, R in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
        getattr(L, f"oracle_feq_{suf}").argtypes = [P, P, ctypes.c_longlong]
        getattr(L, f"oracle_rho_u_{suf}").argtypes = [P, P, ctypes.c_longlong]
        getattr(L, f"oracle_collide_{suf}").argtypes = [P, ctypes.c_longlong, R, R]
        getattr(L, f"oracle_omega_les_{suf}").argtypes = [P, R, R]
        getattr(L, f"oracle_omega_les_{suf}").restype = R
        getattr(L, f"oracle_rescale_{suf}").argtypes = [P, ctypes.c_longlong, ctypes.c_int]
        getattr(L, f"oracle_bounceback_{suf}").argtypes = [P, ctypes.c_longlong]
        getattr(L, f"oracle_bc_{suf}").argtypes = [ctypes.c_int, ctypes.c_int, P, P]
        getattr(L, f"oracle_local_update_{suf}").argtypes = [P, P, ctypes.POINTER(OracleInfo), P, P, P]
        getattr(L, f"oracle_periodic_{suf}").argtypes = [P, P, ctypes.POINTER(OracleInfo), ctypes.c_int]
        getattr(L, f"oracle_fill_equilibrium_{suf}").argtypes = [P, ctypes.POINTER(OracleInfo), P, P, P, P]
        getattr(L, f"oracle_momentum_exchange_{suf}").argtypes = [P, P, ctypes.POINTER(OracleInfo), P, P, P, P]
        getattr(L, f"oracle_mass_{suf}").argtypes = [P, P, ctypes.POINTER(OracleInfo), P]
        getattr(L, f"oracle_mass_{suf}").restype = ctypes.c_double
        getattr(L, f"oracle_coarse_fine_{suf}").argtypes = [P, P, ctypes.POINTER(OracleInfo), P, P,
                                                            ctypes.POINTER(OracleInfo), P]
        getattr(L, f"oracle_fill_coarse_from_fine_{suf}").argtypes = [P, P, ctypes.POINTER(OracleInfo), P, P,
                                                                      ctypes.POINTER(OracleInfo)]
        getattr(L, f"oracle_esotwist_update_{suf}").argtypes = [P, ctypes.c_longlong, P, P,
                                                                ctypes.POINTER(OracleInfo), P, P, ctypes.c_int]
    L.oracle_stream_shift.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.oracle_ray_hit.argtypes = [ctypes.c_longlong] * 6
    L.oracle_ray_hit_z.argtypes = [P, ctypes.c_float, ctypes.c_float]
    L.oracle_ray_hit_z.restype = ctypes.c_float
    L.oracle_markers_inside_stl.argtypes = [P, ctypes.c_int, P, P, ctypes.c_float, ctypes.c_int, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int, P]
    L.oracle_check_stl_edges.argtypes = [P, ctypes.c_int, P]
    L.oracle_esotwist_tables.argtypes = [P, P]
    for suf in ("f32", "f64"):
        for name in ("to_coarse", "to_fine"):
            getattr(L, f"oracle_eso_{name}_{suf}").argtypes = [P, ctypes.c_longlong, P, ctypes.c_int, P,
                                                                ctypes.c_longlong, P, ctypes.c_int, P, P,
                                                                ctypes.c_longlong]
        getattr(L, f"oracle_eso_momentum_exchange_{suf}").argtypes = [P, ctypes.c_longlong, P, ctypes.c_int, P, P]
        getattr(L, f"oracle_eso_moving_bounceback_{suf}").argtypes = [P, ctypes.c_longlong, P, ctypes.c_int, P,
                                                                      ctypes.c_longlong, P]
        RT = ctypes.c_float if suf == "f32" else ctypes.c_double
        getattr(L, f"oracle_eso_forced_velocity_{suf}").argtypes = [P, ctypes.c_longlong, P, ctypes.c_int, P,
                                                                     ctypes.c_longlong, P, RT, RT]
        getattr(L, f"oracle_eso_forcing_sum_{suf}").argtypes = [P, ctypes.c_longlong, P, ctypes.c_int, P,
                                                                 ctypes.c_longlong, P, P] + [ctypes.c_float] * 6 + [P]
        getattr(L, f"oracle_collide_forced_{suf}").argtypes = [P, RT, RT, P]
        getattr(L, f"oracle_scalar_update_{suf}").argtypes = [P, P, P, ctypes.POINTER(OracleInfo), P, P, P, RT]
        getattr(L, f"oracle_scalar_outlet_x_{suf}").argtypes = [P, P, ctypes.POINTER(OracleInfo), P]
        getattr(L, f"oracle_scalar_collide_{suf}").argtypes = [P, P, ctypes.c_longlong, RT]
        getattr(L, f"oracle_eso_outlet_z_{suf}").argtypes = [P, ctypes.c_longlong, P, ctypes.c_int, P,
                                                             ctypes.c_longlong, P]
    L.oracle_diad_build.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_float, P, P, P, ctypes.c_int, P,
                                    ctypes.c_int]
    L.oracle_diad_build.restype = P
    L.oracle_diad_error.argtypes = [P]
    L.oracle_diad_error.restype = ctypes.c_char_p
    L.oracle_diad_count.argtypes = [P, ctypes.c_int, ctypes.c_int]
    L.oracle_diad_count.restype = ctypes.c_longlong
    L.oracle_diad_get.argtypes = [P, ctypes.c_int, ctypes.c_int, P]
    L.oracle_diad_info.argtypes = [P, ctypes.c_int, P, P]
    L.oracle_diad_free.argtypes = [P]


def _suf(a: np.ndarray) -> str:
    if a.dtype == np.float32:
        return "f32"
    if a.dtype == np.float64:
        return "f64"
    raise TypeError(a.dtype)


def _p(a):
    return None if a is None else a.ctypes.data


def _c(a, dtype):
    return None if a is None else np.ascontiguousarray(a, dtype=dtype)


# ------------------------------------------------------------------ per-cell
def feq(rho_u: np.ndarray) -> np.ndarray:
    rho_u = np.ascontiguousarray(rho_u)
    out = np.empty(rho_u.shape[:-1] + (27,), rho_u.dtype)
    getattr(lib(), f"oracle_feq_{_suf(rho_u)}")(rho_u.ctypes.data, out.ctypes.data, out.size // 27)
    return out


def rho_u(f: np.ndarray) -> np.ndarray:
    f = np.ascontiguousarray(f)
    out = np.empty(f.shape[:-1] + (4,), f.dtype)
    getattr(lib(), f"oracle_rho_u_{_suf(f)}")(f.ctypes.data, out.ctypes.data, f.size // 27)
    return out


def collide(f: np.ndarray, nu: float, cs: float) -> np.ndarray:
    f = np.array(f, copy=True, order="C")
    getattr(lib(), f"oracle_collide_{_suf(f)}")(f.ctypes.data, f.size // 27, nu, cs)
    return f


def omega_les(f: np.ndarray, nu: float, cs: float) -> float:
    f = np.ascontiguousarray(f)
    return getattr(lib(), f"oracle_omega_les_{_suf(f)}")(f.ctypes.data, nu, cs)


def rescale(f: np.ndarray, coarse_to_fine: bool) -> np.ndarray:
    f = np.array(f, copy=True, order="C")
    getattr(lib(), f"oracle_rescale_{_suf(f)}")(f.ctypes.data, f.size // 27, int(coarse_to_fine))
    return f


def bounceback(f: np.ndarray) -> np.ndarray:
    f = np.array(f, copy=True, order="C")
    getattr(lib(), f"oracle_bounceback_{_suf(f)}")(f.ctypes.data, f.size // 27)
    return f


def bc(which: int, code: int, f: np.ndarray, rho_u_in: np.ndarray):
    """which: 0 applyMBBC, 1 restoreRhoUxUyUz, 2 restoreUxUyUz, 3 restoreRho."""
    f = np.array(f, copy=True, order="C")
    ru = np.array(rho_u_in, dtype=f.dtype, copy=True)
    rc = getattr(lib(), f"oracle_bc_{_suf(f)}")(which, code, f.ctypes.data, ru.ctypes.data)
    if rc != 0:
        raise ValueError(f"oracle_bc rc={rc}")
    return f, ru


def normal_code(nx, ny, nz):
    return (nx + 5) * 100 + (ny + 5) * 10 + (nz + 5)


# ------------------------------------------------------------------ grid level
class Grid:
    """One PS grid: f (27, N) SoA, shifter[27], flags[N], BC value table (4 x 4)."""

    def __init__(self, nx, ny, nz, dtype=np.float64, nu=0.0, cs=0.0, flags=None, bcvals=None, sub=None,
                 cs_cell=None):
        self.nx, self.ny, self.nz = nx, ny, nz
        self.n = nx * ny * nz
        self.dtype = np.dtype(dtype)
        self.f = np.zeros((27, self.n), self.dtype)
        self.shift = np.zeros(27, np.int32)
        self.flags = None if flags is None else np.ascontiguousarray(flags, np.uint8).reshape(-1)
        self.bcvals = np.zeros((4, 4), self.dtype) if bcvals is None else np.array(bcvals, self.dtype).reshape(4, 4)
        self.cs_cell = _c(cs_cell, self.dtype)
        self.info = make_info(nx, ny, nz, nu, cs, sub)

    @property
    def suf(self):
        return "f32" if self.dtype == np.float32 else "f64"

    def fill_equilibrium(self, rho=None, ux=None, uy=None, uz=None):
        args = [_c(a, self.dtype) for a in (rho, ux, uy, uz)]
        getattr(lib(), f"oracle_fill_equilibrium_{self.suf}")(self.f.ctypes.data, ctypes.byref(self.info),
                                                             *[_p(a) for a in args])

    def stream(self):
        lib().oracle_stream_shift(self.shift.ctypes.data, self.nx, self.ny, self.nz)

    def periodic(self, axis):
        getattr(lib(), f"oracle_periodic_{self.suf}")(self.f.ctypes.data, self.shift.ctypes.data,
                                                     ctypes.byref(self.info), axis)

    def scalar_update(self, T, tau_t, given_t=None):
        """applyLocalCellUpdate(Grid, ScalarTransport) on raw arrays (in place on
        self.f and T, the scalar populations in the same shifted layout)."""
        g = None if given_t is None else np.ascontiguousarray(given_t, self.f.dtype)
        getattr(lib(), f"oracle_scalar_update_{self.suf}")(self.f.ctypes.data, T.ctypes.data,
                                                          self.shift.ctypes.data, ctypes.byref(self.info),
                                                          _p(self.flags), self.bcvals.ctypes.data, _p(g), tau_t)

    def scalar_outlet_x(self, T):
        """applyZeroGradientOutletScalarTransportBC on the raw scalar array T (in place)."""
        getattr(lib(), f"oracle_scalar_outlet_x_{self.suf}")(T.ctypes.data, self.shift.ctypes.data,
                                                            ctypes.byref(self.info), _p(self.flags))

    def local_update(self):
        getattr(lib(), f"oracle_local_update_{self.suf}")(self.f.ctypes.data, self.shift.ctypes.data,
                                                         ctypes.byref(self.info), _p(self.flags),
                                                         self.bcvals.ctypes.data, _p(self.cs_cell))

    def step(self, periodic_x=False, periodic_y=False):
        self.stream()
        if periodic_x:
            self.periodic(0)
        if periodic_y:
            self.periodic(1)
        self.local_update()

    def logical_f(self) -> np.ndarray:
        """(27, N) in logical cell order: logical[q, cell] = f[q, (cell+shift[q]) % N]."""
        out = np.empty_like(self.f)
        idx = np.arange(self.n, dtype=np.int64)
        for q in range(27):
            out[q] = self.f[q, (idx + int(self.shift[q])) % self.n]
        return out

    def momentum_exchange(self, lo, hi):
        out = np.zeros(3, np.float64)
        lo = np.asarray(lo, np.int32); hi = np.asarray(hi, np.int32)
        getattr(lib(), f"oracle_momentum_exchange_{self.suf}")(self.f.ctypes.data, self.shift.ctypes.data,
                                                              ctypes.byref(self.info), _p(self.flags),
                                                              lo.ctypes.data, hi.ctypes.data, out.ctypes.data)
        return out

    def mass(self):
        return getattr(lib(), f"oracle_mass_{self.suf}")(self.f.ctypes.data, self.shift.ctypes.data,
                                                        ctypes.byref(self.info), _p(self.flags))


def coarse_fine_comm(coarse: Grid, fine: Grid):
    getattr(lib(), f"oracle_coarse_fine_{coarse.suf}")(coarse.f.ctypes.data, coarse.shift.ctypes.data,
                                                      ctypes.byref(coarse.info), fine.f.ctypes.data,
                                                      fine.shift.ctypes.data, ctypes.byref(fine.info),
                                                      _p(fine.flags))


def fill_coarse_from_fine(coarse: Grid, fine: Grid):
    getattr(lib(), f"oracle_fill_coarse_from_fine_{coarse.suf}")(coarse.f.ctypes.data, coarse.shift.ctypes.data,
                                                                ctypes.byref(coarse.info), fine.f.ctypes.data,
                                                                fine.shift.ctypes.data, ctypes.byref(fine.info))


def esotwist_update(f, nbr7, ijk3, info, flags, bcvals, flipper):
    getattr(lib(), f"oracle_esotwist_update_{_suf(f)}")(f.ctypes.data, f.shape[1], nbr7.ctypes.data,
                                                        ijk3.ctypes.data, ctypes.byref(info), _p(flags),
                                                        bcvals.ctypes.data, int(flipper))


def stream_shift(shift, nx, ny, nz):
    shift = np.array(shift, np.int32, copy=True)
    lib().oracle_stream_shift(shift.ctypes.data, nx, ny, nz)
    return shift


def velocity_set():
    """(cx, cy, cz), 27 ints each, in the reference ordering."""
    c = [np.zeros(27, np.int32) for _ in range(3)]
    lib().oracle_velocity_set(*[a.ctypes.data for a in c])
    return c


def ray_hit(ax, ay, bx, by, cx, cy) -> bool:
    return bool(lib().oracle_ray_hit(ax, ay, bx, by, cx, cy))


def ray_hit_z(tri9, rx, ry) -> float:
    t = np.ascontiguousarray(tri9, np.float32)
    return lib().oracle_ray_hit_z(t.ctypes.data, rx, ry)


def markers_inside_stl(tri, stl_origin, grid_origin, res, nx, ny, nz, inside_value=1):
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    so = np.ascontiguousarray(stl_origin, np.float32)
    go = np.ascontiguousarray(grid_origin, np.float32)
    out = np.zeros(nx * ny * nz, np.uint8)
    lib().oracle_markers_inside_stl(tri.ctypes.data, tri.shape[0], so.ctypes.data, go.ctypes.data, res, nx, ny, nz,
                                    int(inside_value), out.ctypes.data)
    return out


def check_stl_edges(tri):
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    counts = np.zeros(3 * tri.shape[0], np.int32)
    bad = lib().oracle_check_stl_edges(tri.ctypes.data, tri.shape[0], counts.ctypes.data)
    return bad, counts.reshape(-1, 3)


def esotwist_tables():
    r = np.zeros(27, np.int32)
    w = np.zeros(27, np.int32)
    lib().oracle_esotwist_tables(r.ctypes.data, w.ctypes.data)
    return r, w


def eso_to_coarse(fc, nbc7, flip_c, ff, nbf7, flip_f, write, read):
    """writeToCoarseGridInterface(DIAD) on raw esotwist arrays f[q, cell] (in place on fc)."""
    w = np.ascontiguousarray(write, np.int32)
    r = np.ascontiguousarray(read, np.int32)
    getattr(lib(), f"oracle_eso_to_coarse_{_suf(fc)}")(fc.ctypes.data, fc.shape[1], nbc7.ctypes.data, int(flip_c),
                                                       ff.ctypes.data, ff.shape[1], nbf7.ctypes.data, int(flip_f),
                                                       w.ctypes.data, r.ctypes.data, w.size)


def eso_to_fine(fc, nbc7, flip_c, ff, nbf7, flip_f, write, read):
    """writeToFineGridInterface(DIAD) on raw esotwist arrays (in place on ff)."""
    w = np.ascontiguousarray(write, np.int32)
    r = np.ascontiguousarray(read, np.int32)
    getattr(lib(), f"oracle_eso_to_fine_{_suf(fc)}")(fc.ctypes.data, fc.shape[1], nbc7.ctypes.data, int(flip_c),
                                                     ff.ctypes.data, ff.shape[1], nbf7.ctypes.data, int(flip_f),
                                                     w.ctypes.data, r.ctypes.data, w.size)


def scalar_collide(f, T, tau_t):
    """applyScalarTransportCollision per cell: f, T as (..., 27); returns the new T."""
    f = np.ascontiguousarray(f)
    T = np.array(T, dtype=f.dtype, copy=True, order="C")
    getattr(lib(), f"oracle_scalar_collide_{_suf(f)}")(f.ctypes.data, T.ctypes.data, f.size // 27, tau_t)
    return T


def eso_outlet_z(f, nbr7, flipper, outlet, kmin):
    """applyNonReflectiveOutletZ on a raw esotwist array (in place)."""
    o = np.ascontiguousarray(outlet, np.int32)
    k = np.ascontiguousarray(kmin, np.int32)
    getattr(lib(), f"oracle_eso_outlet_z_{_suf(f)}")(f.ctypes.data, f.shape[1], nbr7.ctypes.data, int(flipper),
                                                     o.ctypes.data, o.size, k.ctypes.data)


def eso_moving_bounceback(f, nbr7, flipper, cells, u3):
    """applyMovingBounceback on listed cells of a raw esotwist array (in place)."""
    c = np.ascontiguousarray(cells, np.int32)
    u = np.ascontiguousarray(u3, np.float32).reshape(-1)
    getattr(lib(), f"oracle_eso_moving_bounceback_{_suf(f)}")(f.ctypes.data, f.shape[1], nbr7.ctypes.data,
                                                              int(flipper), c.ctypes.data, c.size, u.ctypes.data)


def collide_forced(f, nu, cs, g):
    """The forcing collision of one cell (27 values) with force g (3 values)."""
    f = np.array(f, copy=True, order="C")
    g = np.ascontiguousarray(g, f.dtype)
    getattr(lib(), f"oracle_collide_forced_{_suf(f)}")(f.ctypes.data, nu, cs, g.ctypes.data)
    return f


def eso_forced_velocity(f, nbr7, flipper, cells, ut3, nu, cs):
    """Forced-velocity cells (read, force towards the targets ut3, forcing
    collision, write) on a raw esotwist array, in place."""
    c = np.ascontiguousarray(cells, np.int32)
    u = np.ascontiguousarray(ut3, np.float32).reshape(-1)
    getattr(lib(), f"oracle_eso_forced_velocity_{_suf(f)}")(f.ctypes.data, f.shape[1], nbr7.ctypes.data,
                                                            int(flipper), c.ctypes.data, c.size, u.ctypes.data,
                                                            nu, cs)


def eso_forcing_sum(f, nbr7, flipper, cells, ut3, ijk, info, rho_phys):
    """(Fx, Fy, Fz) in N and the torque about the origin (Tx, Ty, Tz) in N m
    of the forced-velocity cells.  info: dict(res, ox, oy, oz, dt)."""
    c = np.ascontiguousarray(cells, np.int32)
    u = np.ascontiguousarray(ut3, np.float32).reshape(-1)
    ijk = np.ascontiguousarray(ijk, np.int32)
    out = np.zeros(6, np.float64)
    getattr(lib(), f"oracle_eso_forcing_sum_{_suf(f)}")(f.ctypes.data, f.shape[1], nbr7.ctypes.data, int(flipper),
                                                        c.ctypes.data, c.size, u.ctypes.data, ijk.ctypes.data,
                                                        info["res"], info["ox"], info["oy"], info["oz"],
                                                        info["dt"], rho_phys, out.ctypes.data)
    return out


def eso_momentum_exchange(f, nbr7, flipper, flags):
    out = np.zeros(3, np.float64)
    fl = np.ascontiguousarray(flags, np.uint8)
    getattr(lib(), f"oracle_eso_momentum_exchange_{_suf(f)}")(f.ctypes.data, f.shape[1], nbr7.ctypes.data,
                                                              int(flipper), fl.ctypes.data, out.ctypes.data)
    return out


# ------------------------------------------------------------ DIAD builder
# Marker rule record (same layout as the library's psl_marker_rule).
RULE_DTYPE = np.dtype([("what", "<i4"), ("value", "<i4"), ("level_mask", "<u4"), ("shape", "<i4"),
                       ("p", "<f4", (8,)), ("idx", "<i4", (6,))])
DIAD_ARRAYS = ("ijk", "bounceback", "interface", "nbr", "c2f_read", "c2f_write", "f2c_write", "f2c_read", "outlet")


def diad_build(base, levels, span, nu_phys, stls=(), rules=None):
    """buildDIADGrids on the CPU.  base: dict(res, ox, oy, oz, nu, dt, nx, ny, nz);
    stls: [(tri (n x 9), origin (3,))]; rules: RULE_DTYPE array.  Returns one
    dict per level: the arrays named in DIAD_ARRAYS (ijk as 3 x n, nbr as 10 x n)
    and the level's info."""
    f6 = np.array([base["res"], base["ox"], base["oy"], base["oz"], base.get("nu", 0.0), base.get("dt", 1.0)],
                  np.float32)
    i3 = np.array([base["nx"], base["ny"], base["nz"]], np.int32)
    tris = [np.ascontiguousarray(t, np.float32).reshape(-1, 9) for t, _ in stls]
    alltri = np.ascontiguousarray(np.concatenate(tris) if tris else np.zeros((0, 9), np.float32))
    counts = np.array([t.shape[0] for t in tris] or [0], np.int32)
    origins = np.array([o for _, o in stls] or [(0, 0, 0)], np.float32).reshape(-1)
    r = np.zeros(0, RULE_DTYPE) if rules is None else np.ascontiguousarray(rules, RULE_DTYPE)
    L = lib()
    h = L.oracle_diad_build(f6.ctypes.data, i3.ctypes.data, levels, span, nu_phys, alltri.ctypes.data,
                            counts.ctypes.data, origins.ctypes.data, len(tris), r.ctypes.data if r.size else None,
                            r.size)
    try:
        err = L.oracle_diad_error(h).decode()
        if err:
            raise RuntimeError(err)
        out = []
        for lev in range(levels):
            fi = np.zeros(6, np.float32)
            ii = np.zeros(4, np.int32)
            L.oracle_diad_info(h, lev, fi.ctypes.data, ii.ctypes.data)
            d = {"res": fi[0], "ox": fi[1], "oy": fi[2], "oz": fi[3], "nu": fi[4], "dt": fi[5],
                 "nx": int(ii[0]), "ny": int(ii[1]), "nz": int(ii[2]), "n": int(ii[3])}
            for w, name in enumerate(DIAD_ARRAYS):
                cnt = L.oracle_diad_count(h, lev, w)
                a = np.zeros(cnt, np.uint8 if name in ("bounceback", "interface") else np.int32)
                if cnt:
                    L.oracle_diad_get(h, lev, w, a.ctypes.data)
                if name == "ijk":
                    a = a.reshape(3, -1)
                elif name == "nbr":
                    a = a.reshape(10, -1)
                d[name] = a
            out.append(d)
        return out
    finally:
        L.oracle_diad_free(h)


def build_flags() -> str:
    """The compiler flags the oracle library was built with."""
    return lib().oracle_build_flags().decode()
