"""MI355X-native D3Q27 periodic-shift LBM step (HIP, gfx950) behind a C ABI.

Python mirror of the reference's operator interface (stepanmuller/PeriodicShiftLBM
include/*.h): the same names, argument meanings and call order, implemented
over libpsl_amd.so (include/psl.h) through ctypes.  Device memory, streams
and kernels belong to the HIP library; numpy only carries host data in and out.

Reference                                   here
-----------------------------------------   -------------------------------------------
InfoStruct (include/types.h:73-75)          InfoStruct
GridStruct (include/types.h:77)             GridStruct (owns a psl_grid handle)
getMarkers / bouncebackMarkerArray          GridStruct.set_markers(getMarkers) -> flag bytes
getGivenRhoUxUyUz                           GridStruct.set_given_values(...)
getSmagorinskyConstant                      GridStruct.set_smagorinsky(cs or array)
fillEquilibriumFromFunction                 fillEquilibriumFromFunction(grid, getInitialRhoUxUyUz)
applyStreaming / applyLocalCellUpdate       applyStreaming / applyLocalCellUpdate
applyBCPeriodicX / applyBCPeriodicY         applyBCPeriodicX / applyBCPeriodicY
applyCoarseFineGridCommunication            applyCoarseFineGridCommunication
fillCoarseGridFromFine                      fillCoarseGridFromFine
updateGrid recursion                        updateGrid(grids, steps)
readSTL / checkSTLEdges                     readSTL / checkSTLEdges
applyMarkersInsideSTL, bool array ops       applyMarkersInsideSTL, multiplyBoolArrays, ...
getSphereDrag / getLift                     momentumExchange (FP64)

There is no CPU fallback: without the built library and a HIP device every
compute call raises.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass, field
from typing import Callable, Optional, Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpsl_amd.so")

F32, F16DEV, BF16DEV = 0, 1, 2
TUNE_NT_LOAD, TUNE_NT_STORE, TUNE_XCD_REMAP, TUNE_FUSED_BOUNDARY = 1, 2, 4, 8
TUNE_VARIANTS = 16
STORAGE_NAMES = {"f32": F32, "f16": F16DEV, "bf16": BF16DEV}

FLUID, BOUNCEBACK, GIVEN_RHO, GIVEN_U, GIVEN_NONE, GIVEN_RHOU, GHOST, COVERED = range(8)
PERIODIC_X, PERIODIC_Y, PERIODIC_Z = 8, 16, 32

CX = np.array([0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, -1, 1, 0, 0, -1, 1, 0, 0, -1, 1, -1, 1, 1, -1, -1, 1])
CY = np.array([0, 0, 0, 0, 0, -1, 1, 0, 0, 0, 0, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1, -1, 1, -1, 1])
CZ = np.array([0, 0, 0, -1, 1, 0, 0, -1, 1, 1, -1, 0, 0, -1, 1, 0, 0, 1, -1, -1, 1, 1, -1, -1, 1, -1, 1])
W = np.array([8 / 27] + [2 / 27] * 6 + [1 / 54] * 12 + [1 / 216] * 8)


class PslError(RuntimeError):
    pass


class _Info(ctypes.Structure):
    _fields_ = [("gridID", ctypes.c_float), ("res", ctypes.c_float),
                ("cellCountX", ctypes.c_int), ("cellCountY", ctypes.c_int), ("cellCountZ", ctypes.c_int),
                ("cellCount", ctypes.c_int),
                ("ox", ctypes.c_float), ("oy", ctypes.c_float), ("oz", ctypes.c_float),
                ("nu", ctypes.c_float), ("dtPhys", ctypes.c_float),
                ("iSubgridStart", ctypes.c_int), ("jSubgridStart", ctypes.c_int), ("kSubgridStart", ctypes.c_int),
                ("iSubgridEnd", ctypes.c_int), ("jSubgridEnd", ctypes.c_int), ("kSubgridEnd", ctypes.c_int),
                ("cellCountFineToCoarse", ctypes.c_int), ("cellCountCoarseToFine", ctypes.c_int),
                ("pRegulator", ctypes.c_float), ("iRegulator", ctypes.c_float)]


@dataclass
class InfoStruct:
    """include/types.h:73-75."""
    gridID: float = 0.0
    res: float = 1.0
    cellCountX: int = 0
    cellCountY: int = 0
    cellCountZ: int = 0
    cellCount: int = 0
    ox: float = 0.0
    oy: float = 0.0
    oz: float = 0.0
    nu: float = 1.0
    dtPhys: float = 1.0
    iSubgridStart: int = 0
    jSubgridStart: int = 0
    kSubgridStart: int = 0
    iSubgridEnd: int = 0
    jSubgridEnd: int = 0
    kSubgridEnd: int = 0
    cellCountFineToCoarse: int = 0
    cellCountCoarseToFine: int = 0
    pRegulator: float = 0.0
    iRegulator: float = 0.0

    def to_c(self) -> _Info:
        c = _Info()
        for name, _ in _Info._fields_:
            setattr(c, name, getattr(self, name))
        c.cellCount = self.cellCountX * self.cellCountY * self.cellCountZ
        return c

    @staticmethod
    def from_c(c: _Info) -> "InfoStruct":
        return InfoStruct(**{name: getattr(c, name) for name, _ in _Info._fields_})


# ---------------------------------------------------------------- library
_lib = None

# (name, restype, argtypes); every symbol declared in include/psl.h
_P = ctypes.c_void_p
_I = ctypes.c_int
_SIGS = [
    ("psl_last_error", ctypes.c_char_p, []),
    ("psl_abi_version", _I, []),
    ("psl_device_count", _I, [_P]),
    ("psl_grid_create", _I, [_P, _I, _P]),
    ("psl_grid_destroy", _I, [_P]),
    ("psl_grid_get_info", _I, [_P, _P]),
    ("psl_grid_set_info", _I, [_P, _P]),
    ("psl_grid_set_stream", _I, [_P, _P]),
    ("psl_grid_set_flags", _I, [_P, _P]),
    ("psl_grid_get_flags", _I, [_P, _P]),
    ("psl_grid_set_bc_values", _I, [_P, _P]),
    ("psl_grid_set_smagorinsky", _I, [_P, ctypes.c_float, _P]),
    ("psl_grid_set_periodic_copies", _I, [_P, _I]),
    ("psl_fill_equilibrium", _I, [_P, _P, _P, _P, _P]),
    ("psl_apply_streaming", _I, [_P]),
    ("psl_apply_local_cell_update", _I, [_P]),
    ("psl_apply_bc_periodic_x", _I, [_P]),
    ("psl_apply_bc_periodic_y", _I, [_P]),
    ("psl_coarse_fine_communication", _I, [_P, _P]),
    ("psl_fill_coarse_from_fine", _I, [_P, _P]),
    ("psl_update_grid", _I, [_P, _I, _I]),
    ("psl_get_shifter", _I, [_P, _P]),
    ("psl_set_shifter", _I, [_P, _P]),
    ("psl_get_f", _I, [_P, _P]),
    ("psl_set_f", _I, [_P, _P]),
    ("psl_get_macroscopic", _I, [_P, _P, _P, _P, _P]),
    ("psl_grid_device_f", _I, [_P, _P, _P]),
    ("psl_reduce_momentum_exchange", _I, [_P, _P, _P, _P]),
    ("psl_reduce_momentum_exchange_noalloc", _I, [_P, _P, _P, _P]),
    ("psl_reduce_mass", _I, [_P, _P]),
    ("psl_stl_read", _I, [ctypes.c_char_p, _P]),
    ("psl_stl_from_triangles", _I, [_P, _I, _P]),
    ("psl_stl_destroy", _I, [_P]),
    ("psl_stl_info", _I, [_P, _P, _P]),
    ("psl_stl_set_origin", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float]),
    ("psl_stl_get_triangles", _I, [_P, _P]),
    ("psl_stl_rotate", _I, [_P, _I, ctypes.c_float]),
    ("psl_diad_markers_inside_stl", _I, [_P, _P, _P]),
    ("psl_stl_check_edges", _I, [_P, _P, _P]),
    ("psl_mask_create", _I, [ctypes.c_int64, _I, _P]),
    ("psl_mask_destroy", _I, [_P]),
    ("psl_mask_get", _I, [_P, _P]),
    ("psl_mask_set", _I, [_P, _P]),
    ("psl_mask_and", _I, [_P, _P, _P]),
    ("psl_mask_or", _I, [_P, _P, _P]),
    ("psl_mask_not", _I, [_P]),
    ("psl_markers_inside_stl", _I, [_P, _P, _I, _P]),
    ("psl_grid_apply_bounceback_mask", _I, [_P, _P]),
    ("psl_device_synchronize", _I, []),
    ("psl_grid_profile_enable", _I, [_P, _I]),
    ("psl_grid_profile_read", _I, [_P, _P, _P]),
    ("psl_grid_cell_census", _I, [_P, _P, _P, _P]),
    ("psl_grid_set_tuning", _I, [_P, _I]),
    ("psl_grid_get_tuning", _I, [_P, _P]),
    ("psl_grid_set_occupancy", _I, [_P, _I]),
    ("psl_grid_get_occupancy", _I, [_P, _P]),
    ("psl_grid_autotune", _I, [_P, _I, _P, _P]),
    ("psl_device_copy", _I, [_P, _P, ctypes.c_size_t]),
    ("psl_device_alloc", _I, [ctypes.c_size_t, _P]),
    ("psl_device_free", _I, [_P]),
    ("psl_batch_create", _I, [_P, _I, _P]),
    ("psl_batch_destroy", _I, [_P]),
    ("psl_batch_step", _I, [_P, _I]),
    ("psl_batch_set_steps_per_graph", _I, [_P, _I]),
    ("psl_batch_get_steps_per_graph", _I, [_P, _P]),
    ("psl_batch_set_tuning", _I, [_P, _I, _I]),
    ("psl_batch_get_tuning", _I, [_P, _P, _P]),
    ("psl_batch_autotune", _I, [_P, _I, _P, _P, _P]),
    ("psl_batch_set_layout", _I, [_P, _I, _I]),
    ("psl_batch_get_layout", _I, [_P, _P, _P]),
    ("psl_batch_autotune_layouts", _I, [_P, _I, _P, _P]),
    ("psl_batch_census", _I, [_P, _P, _P, _P]),
    ("psl_device_upload", _I, [_P, _P, ctypes.c_size_t]),
    ("psl_device_download", _I, [_P, _P, ctypes.c_size_t]),
    ("psl_mlp_create", _I, [_P]),
    ("psl_mlp_destroy", _I, [_P]),
    ("psl_mlp_set_params", _I, [_P, _P, _P, _P]),
    ("psl_mlp_get_params", _I, [_P, _P, _P, _P]),
    ("psl_mlp_loss_grad", _I, [_P, _P, _P, _I, _P, _P]),
    ("psl_mlp_train", _I, [_P, _P, _P, _I, _I, _I, ctypes.c_float, ctypes.c_uint64, _P]),
    ("psl_mlp_eval", _I, [_P, _P, _I, _P]),
    ("psl_les_filter", _I, [_P, _P, _P, _P, _I, _I, _I, _I, _P]),
    ("psl_les_targets", _I, [_P, _I, ctypes.c_float, _P, _P]),
    ("psl_grid_set_closure", _I, [_P, _P]),
    ("psl_grid_set_closure_every", _I, [_P, _I]),
    ("psl_grid_get_closure_y", _I, [_P, _P]),
    ("psl_get_macroscopic_device", _I, [_P, _P]),
    ("psl_grid_time_update_kernel", _I, [_P, _I, _P, _P]),
    ("psl_grid_time_update_kernel_streaming", _I, [_P, _I, _P, _P]),
    ("psl_probe_ps_pattern", _I, [_I, _I, _I, _I, _I, _P]),
    ("psl_probe_ps_pattern_capped", _I, [_I, _I, _I, _I, _I, _I, _P]),
    ("psl_probe_ps_access", _I, [_I, _I, _I, _I, _I, _I, _I, _I, _I, _P]),
    ("psl_diad_create", _I, [_P, _I, ctypes.c_int64, _P, _P, _P]),
    ("psl_diad_destroy", _I, [_P]),
    ("psl_diad_set_flags", _I, [_P, _P]),
    ("psl_diad_set_bc_values", _I, [_P, _P]),
    ("psl_diad_set_smagorinsky", _I, [_P, ctypes.c_float]),
    ("psl_diad_set_stream", _I, [_P, _P]),
    ("psl_diad_fill_equilibrium", _I, [_P, _P, _P, _P, _P]),
    ("psl_diad_apply_streaming", _I, [_P]),
    ("psl_diad_apply_local_cell_update", _I, [_P]),
    ("psl_diad_get_flipper", _I, [_P, _P]),
    ("psl_diad_get_f", _I, [_P, _P]),
    ("psl_diad_set_f", _I, [_P, _P]),
    ("psl_diad_dense_box", _I, [_I, _I, _I, _I, _P, _P]),
    ("psl_diad_build", _I, [_P, _I, _I, ctypes.c_float, _P, _I, _P, _I, _P]),
    ("psl_diad_levels_destroy", _I, [_P]),
    ("psl_diad_levels_info", _I, [_P, _I, _P, _P]),
    ("psl_diad_levels_get", _I, [_P, _I, _I, _P]),
    ("psl_diad_levels_build_ms", _I, [_P, _P]),
    ("psl_diad_create_level", _I, [_P, _I, _I, _P]),
    ("psl_diad_set_interface", _I, [_P, _P, _P, _P, ctypes.c_int64, _P, _P, ctypes.c_int64]),
    ("psl_diad_coarse_fine_communication", _I, [_P, _P]),
    ("psl_diad_update_grid", _I, [_P, _I, _I]),
    ("psl_diad_interface_records", _I, [_P, _P, _P]),
    ("psl_diad_state_bytes", _I, [_P, _P]),
    ("psl_diad_save_state", _I, [_P, _P, ctypes.c_int64]),
    ("psl_diad_load_state", _I, [_P, _P, ctypes.c_int64]),
    ("psl_diad_check_health", _I, [_P, _P, _P, _P]),
    ("psl_grid_state_bytes", _I, [_P, _P]),
    ("psl_grid_save_state", _I, [_P, _P, ctypes.c_int64]),
    ("psl_grid_load_state", _I, [_P, _P, ctypes.c_int64]),
    ("psl_grid_check_health", _I, [_P, _P, _P, _P]),
    ("psl_diad_verify_interface_records", _I, [_P, _P]),
    ("psl_diad_reduce_momentum_exchange", _I, [_P, _P]),
    ("psl_diad_set_outlet", _I, [_P, _P, ctypes.c_int64, _P]),
    ("psl_diad_apply_nonreflective_outlet_z", _I, [_P]),
    ("psl_diad_set_moving_bounceback", _I, [_P, _P, ctypes.c_int64, _P]),
    ("psl_diad_set_forced_velocity", _I, [_P, _P, ctypes.c_int64, _P]),
    ("psl_diad_reduce_forcing", _I, [_P, ctypes.c_float, _P]),
    ("psl_section_cut", _I, [_P, _I, _I, ctypes.c_float, ctypes.c_float, _P, _P]),
    ("psl_scalar_section_cut", _I, [_P, _I, _I, ctypes.c_float, ctypes.c_float, _P, _P]),
    ("psl_volume_export", _I, [_P, ctypes.c_float, ctypes.c_float, _P]),
    ("psl_cylinder_cut", _I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P, _P]),
    ("psl_diad_cylinder_cut", _I, [_P, _I, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_int64, _P, _P]),
    ("psl_write_volume_data", _I, [ctypes.c_char_p, _I, _I, _I, _I, _I, _P]),
    ("psl_diad_section_cut", _I, [_P, _I, _I, _I, ctypes.c_float, ctypes.c_float, ctypes.c_int64, _P, _P]),
    ("psl_scalar_enable", _I, [_P, ctypes.c_float]),
    ("psl_scalar_set_given", _I, [_P, _P]),
    ("psl_scalar_get", _I, [_P, _P]),
    ("psl_scalar_set", _I, [_P, _P]),
    ("psl_scalar_apply_local_cell_update", _I, [_P]),
    ("psl_scalar_apply_zero_gradient_outlet_x", _I, [_P]),
    ("psl_diad_flow_report", _I, [_P, _I, _I, _I, _P, _P]),
    ("psl_write_sim_data", _I, [ctypes.c_char_p, _I, _I, _I, _I, _P]),
    ("psl_write_history", _I, [ctypes.c_char_p, _P, _I]),
]


def lib():
    """Load libpsl_amd.so; raises when it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise PslError(f"{LIB_PATH} is missing: build it with `python periodicshiftlbm_amd/build.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in _SIGS:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(rc: int):
    if rc != 0:
        msg = lib().psl_last_error()
        raise PslError(f"psl error {rc}: {msg.decode() if msg else ''}")


def device_count() -> int:
    c = ctypes.c_int(0)
    _check(lib().psl_device_count(ctypes.byref(c)))
    return c.value


def require_device():
    if device_count() < 1:
        raise PslError("no HIP device visible: the MI355X path has no CPU fallback")


def _ptr(a):
    return None if a is None else a.ctypes.data


def _state_bytes(state) -> np.ndarray:
    """A restart state given as a byte array or a file path."""
    if isinstance(state, (str, os.PathLike)):
        return np.fromfile(state, np.uint8)
    return np.ascontiguousarray(state, np.uint8).reshape(-1)


def _health(fn, h):
    bad, lo, hi = ctypes.c_int64(), ctypes.c_double(), ctypes.c_double()
    _check(fn(h, ctypes.byref(bad), ctypes.byref(lo), ctypes.byref(hi)))
    return bad.value, lo.value, hi.value


# ---------------------------------------------------------------- grids
class GridStruct:
    """include/types.h:77 -- owns the device fArray, shifter and flags."""

    def __init__(self, info: InfoStruct, storage: int | str = F32):
        if isinstance(storage, str):
            storage = STORAGE_NAMES[storage]
        require_device()
        self.storage = storage
        self._h = ctypes.c_void_p()
        ci = info.to_c()
        _check(lib().psl_grid_create(ctypes.byref(ci), storage, ctypes.byref(self._h)))
        self.Info = InfoStruct.from_c(ci)

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().psl_grid_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def n(self) -> int:
        return self.Info.cellCountX * self.Info.cellCountY * self.Info.cellCountZ

    def push_info(self):
        ci = self.Info.to_c()
        _check(lib().psl_grid_set_info(self._h, ctypes.byref(ci)))

    # -- markers / boundary data (host-evaluated driver callbacks)
    def set_flags(self, flags: Optional[np.ndarray]):
        if flags is None:
            _check(lib().psl_grid_set_flags(self._h, None))
            return
        f = np.ascontiguousarray(flags, np.uint8).reshape(-1)
        if f.size != self.n:
            raise ValueError("flags must hold one byte per cell")
        _check(lib().psl_grid_set_flags(self._h, f.ctypes.data))

    def get_flags(self) -> np.ndarray:
        out = np.empty(self.n, np.uint8)
        _check(lib().psl_grid_get_flags(self._h, out.ctypes.data))
        return out

    def ijk(self):
        I = self.Info
        k, j, i = np.meshgrid(np.arange(I.cellCountZ), np.arange(I.cellCountY), np.arange(I.cellCountX), indexing="ij")
        return i.reshape(-1), j.reshape(-1), k.reshape(-1)

    def set_markers(self, getMarkers: Callable, bounceback: Optional[np.ndarray] = None):
        """Evaluate a vectorised getMarkers(i, j, k, bounceback, Info) -> flag bytes."""
        i, j, k = self.ijk()
        bb = np.zeros(self.n, bool) if bounceback is None else np.asarray(bounceback, bool).reshape(-1)
        self.set_flags(np.asarray(getMarkers(i, j, k, bb, self.Info), np.uint8))

    def set_given_values(self, values):
        """getGivenRhoUxUyUz table: up to 4 slots of (rho, ux, uy, uz)."""
        v = np.zeros((4, 4), np.float32)
        a = np.asarray(values, np.float32).reshape(-1, 4)
        v[: a.shape[0]] = a
        _check(lib().psl_grid_set_bc_values(self._h, v.ctypes.data))

    def set_smagorinsky(self, cs):
        if np.isscalar(cs):
            _check(lib().psl_grid_set_smagorinsky(self._h, float(cs), None))
        else:
            a = np.ascontiguousarray(cs, np.float32).reshape(-1)
            _check(lib().psl_grid_set_smagorinsky(self._h, 0.0, a.ctypes.data))

    def set_periodic_copies(self, x: bool = False, y: bool = False):
        _check(lib().psl_grid_set_periodic_copies(self._h, int(x) | (int(y) << 1)))

    # -- state
    @property
    def shifter(self) -> np.ndarray:
        out = np.zeros(27, np.int32)
        _check(lib().psl_get_shifter(self._h, out.ctypes.data))
        return out

    @shifter.setter
    def shifter(self, v):
        a = np.ascontiguousarray(v, np.int32)
        _check(lib().psl_set_shifter(self._h, a.ctypes.data))

    def get_f(self) -> np.ndarray:
        """(27, N) FP32 in logical cell order."""
        out = np.empty((27, self.n), np.float32)
        _check(lib().psl_get_f(self._h, out.ctypes.data))
        return out

    def set_f(self, f: np.ndarray):
        a = np.ascontiguousarray(f, np.float32)
        if a.shape != (27, self.n):
            raise ValueError("f must have shape (27, N)")
        _check(lib().psl_set_f(self._h, a.ctypes.data))

    def macroscopic(self):
        n = self.n
        r, a, b, c = (np.empty(n, np.float32) for _ in range(4))
        _check(lib().psl_get_macroscopic(self._h, r.ctypes.data, a.ctypes.data, b.ctypes.data, c.ctypes.data))
        return r, a, b, c

    def macroscopic_device(self) -> "DeviceBuffer":
        """rho, ux, uy, uz (4 x N floats) in a new device buffer."""
        out = DeviceBuffer(4 * self.n * 4)
        _check(lib().psl_get_macroscopic_device(self._h, out.ptr))
        return out

    def state(self) -> np.ndarray:
        """The restart state (psl_grid_save_state): header, f as stored, and the
        scalar populations / held closure output where the grid has them."""
        n = ctypes.c_int64()
        _check(lib().psl_grid_state_bytes(self._h, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        _check(lib().psl_grid_save_state(self._h, buf.ctypes.data, buf.size))
        return buf

    def save_state(self, path):
        self.state().tofile(path)

    def load_state(self, state):
        """Restore a state from state() or save_state (array or path)."""
        a = _state_bytes(state)
        _check(lib().psl_grid_load_state(self._h, a.ctypes.data, a.size))

    def check_health(self):
        """(cells with a non-finite population or rho <= 0, min rho, max rho of the others)."""
        return _health(lib().psl_grid_check_health, self._h)

    def time_update_kernel(self, launches: int = 50, streaming: bool = False):
        """(mean ms, kernel name) of the dominant update kernel launched back to
        back; with `streaming`, the shifter advances before every launch, so the
        launches see the shifts of consecutive steps as in a run."""
        ms = ctypes.c_double()
        k = ctypes.c_int()
        fn = lib().psl_grid_time_update_kernel_streaming if streaming else lib().psl_grid_time_update_kernel
        _check(fn(self._h, int(launches), ctypes.byref(ms), ctypes.byref(k)))
        return ms.value, ("k_ps_bulk", "k_ps_fused", "k_ps_bulk_closure")[k.value]

    def set_closure(self, mlp):
        """Use a learned closure (MLP) in the bulk update; None restores Smagorinsky."""
        self._closure = mlp  # keep the network alive while the grid uses it
        _check(lib().psl_grid_set_closure(self._h, None if mlp is None else mlp.handle))

    def set_closure_every(self, steps: int):
        """Evaluate the network every `steps` local updates and hold its output
        (FP16 per cell) in between; 1 evaluates it every update."""
        _check(lib().psl_grid_set_closure_every(self._h, int(steps)))

    def closure_y(self) -> np.ndarray:
        """The held network output per cell (after a step with set_closure_every > 1)."""
        out = np.zeros(self.n, np.float32)
        _check(lib().psl_grid_get_closure_y(self._h, out.ctypes.data))
        return out

    def device_f(self):
        p = ctypes.c_void_p()
        nbytes = ctypes.c_size_t()
        _check(lib().psl_grid_device_f(self._h, ctypes.byref(p), ctypes.byref(nbytes)))
        return p.value, nbytes.value

    def set_stream(self, stream_ptr: int):
        _check(lib().psl_grid_set_stream(self._h, ctypes.c_void_p(stream_ptr)))

    def momentum_exchange(self, lo=None, hi=None, noalloc: bool = False) -> np.ndarray:
        """FP64 momentum exchange over the bounceback cells of [lo, hi).  With
        noalloc the library's fixed per-device scratch is used instead of a
        buffer it allocates (boxes up to PSL_NOALLOC_MAX_BOX_CELLS cells)."""
        out = np.zeros(3, np.float64)
        lo_a = None if lo is None else np.ascontiguousarray(lo, np.int32)
        hi_a = None if hi is None else np.ascontiguousarray(hi, np.int32)
        fn = lib().psl_reduce_momentum_exchange_noalloc if noalloc else lib().psl_reduce_momentum_exchange
        _check(fn(self._h, _ptr(lo_a), _ptr(hi_a), out.ctypes.data))
        return out

    def profile(self, enable: bool):
        _check(lib().psl_grid_profile_enable(self._h, int(enable)))

    def profile_read(self):
        t = ctypes.c_double()
        n = ctypes.c_int()
        _check(lib().psl_grid_profile_read(self._h, ctypes.byref(t), ctypes.byref(n)))
        return t.value, n.value

    @property
    def tuning(self) -> int:
        v = ctypes.c_int()
        _check(lib().psl_grid_get_tuning(self._h, ctypes.byref(v)))
        return v.value

    @tuning.setter
    def tuning(self, v: int):
        _check(lib().psl_grid_set_tuning(self._h, int(v)))

    @property
    def occupancy(self) -> int:
        """Cap on the bulk kernel's workgroups per CU (0: none; 2-6)."""
        v = ctypes.c_int()
        _check(lib().psl_grid_get_occupancy(self._h, ctypes.byref(v)))
        return v.value

    @occupancy.setter
    def occupancy(self, v: int):
        _check(lib().psl_grid_set_occupancy(self._h, int(v)))

    def autotune(self, steps: int = 20):
        """Time the local-update variants on this grid and keep the fastest, then
        the bulk kernel's occupancy caps (state advances)."""
        best = ctypes.c_int()
        times = np.zeros(TUNE_VARIANTS, np.float64)
        _check(lib().psl_grid_autotune(self._h, steps, ctypes.byref(best), times.ctypes.data))
        return best.value, times

    def census(self):
        a, b, c = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        _check(lib().psl_grid_cell_census(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c)))
        return a.value, b.value, c.value

    def mass(self) -> float:
        out = ctypes.c_double()
        _check(lib().psl_reduce_mass(self._h, ctypes.byref(out)))
        return out.value


class Batch:
    """Many GridStructs advanced together by replays of captured hipGraphs
    (one of a single step, one of steps_per_graph steps)."""

    def __init__(self, grids: Sequence[GridStruct]):
        self.grids = list(grids)  # keep the grids alive
        arr = (ctypes.c_void_p * len(grids))(*[g.handle.value for g in grids])
        self._h = ctypes.c_void_p()
        _check(lib().psl_batch_create(arr, len(grids), ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().psl_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1):
        _check(lib().psl_batch_step(self._h, n))

    @property
    def steps_per_graph(self) -> int:
        v = ctypes.c_int()
        _check(lib().psl_batch_get_steps_per_graph(self._h, ctypes.byref(v)))
        return v.value

    @steps_per_graph.setter
    def steps_per_graph(self, steps: int):
        _check(lib().psl_batch_set_steps_per_graph(self._h, int(steps)))

    def set_tuning(self, tuning: int, order: int = 0):
        _check(lib().psl_batch_set_tuning(self._h, tuning, order))

    @property
    def tuning(self):
        t, o = ctypes.c_int(), ctypes.c_int()
        _check(lib().psl_batch_get_tuning(self._h, ctypes.byref(t), ctypes.byref(o)))
        return t.value, o.value

    def autotune(self, steps: int = 10):
        t, o = ctypes.c_int(), ctypes.c_int()
        times = np.zeros(2 * TUNE_VARIANTS, np.float64)
        _check(lib().psl_batch_autotune(self._h, steps, ctypes.byref(t), ctypes.byref(o), times.ctypes.data))
        return t.value, o.value, times

    def set_layout(self, cells_per_thread: int = 1, tile_order: int = 0):
        """Bulk tiles of 256 * cells_per_thread cells; tile_order 0: each grid's
        tiles in turn, 1: interleaved across grids."""
        _check(lib().psl_batch_set_layout(self._h, int(cells_per_thread), int(tile_order)))

    @property
    def layout(self):
        c, t = ctypes.c_int(), ctypes.c_int()
        _check(lib().psl_batch_get_layout(self._h, ctypes.byref(c), ctypes.byref(t)))
        return c.value, t.value

    def autotune_layouts(self, steps: int = 10):
        """Variant, grid order, cells per thread and tile order, all searched:
        ((variant, order, cells_per_thread, tile_order), times) with times of
        shape (tile_order, cells_per_thread - 1, order, variant), -1 where not run."""
        best = np.zeros(4, np.int32)
        times = np.zeros(8 * TUNE_VARIANTS, np.float64)
        _check(lib().psl_batch_autotune_layouts(self._h, steps, best.ctypes.data, times.ctypes.data))
        return tuple(int(x) for x in best), times.reshape(2, 2, 2, TUNE_VARIANTS)

    def census(self):
        c, b, n = ctypes.c_int64(), ctypes.c_int(), ctypes.c_int()
        _check(lib().psl_batch_census(self._h, ctypes.byref(c), ctypes.byref(b), ctypes.byref(n)))
        return c.value, b.value, n.value


class DIADGridStruct:
    """include/types.h:166-171 -- esotwist grid over an explicit cell list."""

    def __init__(self, info: InfoStruct, ijk3: np.ndarray, nbr7: np.ndarray, storage: int | str = F32,
                 _handle: ctypes.c_void_p | None = None):
        if isinstance(storage, str):
            storage = STORAGE_NAMES[storage]
        require_device()
        if _handle is not None:
            self._h = _handle
            self.Info = info
            self.n = info.cellCount
            self.storage = storage
            return
        ijk3 = np.ascontiguousarray(ijk3, np.int32)
        nbr7 = np.ascontiguousarray(nbr7, np.int32)
        self.n = ijk3.shape[1]
        self.Info = info
        self.storage = storage
        self._h = ctypes.c_void_p()
        ci = info.to_c()
        _check(lib().psl_diad_create(ctypes.byref(ci), storage, self.n, ijk3.ctypes.data, nbr7.ctypes.data,
                                     ctypes.byref(self._h)))

    @staticmethod
    def dense_box(nx, ny, nz, periodic=(True, True, True)):
        n = nx * ny * nz
        ijk = np.zeros((3, n), np.int32)
        nbr = np.zeros((7, n), np.int32)
        mask = int(periodic[0]) | (int(periodic[1]) << 1) | (int(periodic[2]) << 2)
        _check(lib().psl_diad_dense_box(nx, ny, nz, mask, ijk.ctypes.data, nbr.ctypes.data))
        return ijk, nbr

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().psl_diad_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @staticmethod
    def from_levels(levels: "DIADLevels", level: int, storage: int | str = F32) -> "DIADGridStruct":
        """The grid of one built level (tables stay on the device)."""
        st = STORAGE_NAMES[storage] if isinstance(storage, str) else storage
        info, _ = levels.info(level)
        h = ctypes.c_void_p()
        _check(lib().psl_diad_create_level(levels.handle, level, st, ctypes.byref(h)))
        return DIADGridStruct(info, None, None, st, _handle=h)

    def set_interface(self, fine: "DIADGridStruct", c2f_read, c2f_write, f2c_write, f2c_read):
        a = [np.ascontiguousarray(x, np.int32) for x in (c2f_read, c2f_write, f2c_write, f2c_read)]
        _check(lib().psl_diad_set_interface(self._h, fine.handle, a[0].ctypes.data, a[1].ctypes.data, a[0].size,
                                            a[2].ctypes.data, a[3].ctypes.data, a[2].size))

    def interface_records(self):
        """(record count, fresh) of this level's interface records."""
        n = ctypes.c_int64()
        fresh = ctypes.c_int()
        _check(lib().psl_diad_interface_records(self._h, ctypes.byref(n), ctypes.byref(fresh)))
        return n.value, bool(fresh.value)

    def state(self) -> np.ndarray:
        """The restart state (psl_diad_save_state): header with the parity, f as stored."""
        n = ctypes.c_int64()
        _check(lib().psl_diad_state_bytes(self._h, ctypes.byref(n)))
        buf = np.empty(n.value, np.uint8)
        _check(lib().psl_diad_save_state(self._h, buf.ctypes.data, buf.size))
        return buf

    def save_state(self, path):
        self.state().tofile(path)

    def load_state(self, state):
        a = _state_bytes(state)
        _check(lib().psl_diad_load_state(self._h, a.ctypes.data, a.size))

    def check_health(self):
        """(cells with a non-finite population or rho <= 0, min rho, max rho of the others)."""
        return _health(lib().psl_diad_check_health, self._h)

    def verify_interface_records(self) -> int:
        """Records that differ from the populations their cells hold now."""
        n = ctypes.c_int64()
        _check(lib().psl_diad_verify_interface_records(self._h, ctypes.byref(n)))
        return n.value

    def set_moving_bounceback(self, cells, u3):
        """Moving-wall cells (flagged BOUNCEBACK) and their velocities (n x 3)."""
        c = np.ascontiguousarray(cells, np.int32).reshape(-1)
        u = np.ascontiguousarray(u3, np.float32).reshape(-1)
        if u.size != 3 * c.size:
            raise ValueError(f"{c.size} moving-wall cells need {3 * c.size} velocity components, got {u.size}")
        _check(lib().psl_diad_set_moving_bounceback(self._h, c.ctypes.data, c.size, u.ctypes.data))

    def set_forced_velocity(self, cells, ut3):
        """Forced-velocity cells (flagged BOUNCEBACK) and their target
        velocities (n x 3, lattice units)."""
        c = np.ascontiguousarray(cells, np.int32).reshape(-1)
        u = np.ascontiguousarray(ut3, np.float32).reshape(-1)
        if u.size != 3 * c.size:
            raise ValueError(f"{c.size} forced cells need {3 * c.size} velocity components, got {u.size}")
        _check(lib().psl_diad_set_forced_velocity(self._h, c.ctypes.data, c.size, u.ctypes.data))

    def forcing(self, rho_nominal_phys: float) -> np.ndarray:
        """Force (N) and torque about the origin (N m) of the forced cells:
        (Fx, Fy, Fz, Tx, Ty, Tz) in FP64; Tz is getTorque's value."""
        out = np.zeros(6, np.float64)
        _check(lib().psl_diad_reduce_forcing(self._h, rho_nominal_phys, out.ctypes.data))
        return out

    def markers_inside_stl(self, stl: "STLStruct") -> np.ndarray:
        """ApplyMarkersInsideSTL on this grid's cells: 1 inside, 0 outside."""
        out = np.zeros(self.n, np.uint8)
        _check(lib().psl_diad_markers_inside_stl(self._h, stl.handle, out.ctypes.data))
        return out

    def set_outlet(self, outlet_cells, kmin):
        """outletCellArray and kMinNbrArray (n entries)."""
        o = np.ascontiguousarray(outlet_cells, np.int32)
        k = np.ascontiguousarray(kmin, np.int32)
        _check(lib().psl_diad_set_outlet(self._h, o.ctypes.data, o.size, k.ctypes.data))

    def momentum_exchange(self) -> np.ndarray:
        """getSphereDrag(DIADGridStruct&), all three components in FP64."""
        out = np.zeros(3, np.float64)
        _check(lib().psl_diad_reduce_momentum_exchange(self._h, out.ctypes.data))
        return out

    def set_flags(self, flags):
        a = None if flags is None else np.ascontiguousarray(flags, np.uint8)
        _check(lib().psl_diad_set_flags(self._h, _ptr(a)))

    def set_given_values(self, values):
        v = np.zeros((4, 4), np.float32)
        a = np.asarray(values, np.float32).reshape(-1, 4)
        v[: a.shape[0]] = a
        _check(lib().psl_diad_set_bc_values(self._h, v.ctypes.data))

    def set_smagorinsky(self, cs: float):
        _check(lib().psl_diad_set_smagorinsky(self._h, float(cs)))

    def fill_equilibrium(self, rho=None, ux=None, uy=None, uz=None):
        arrs = [None if v is None else np.ascontiguousarray(np.broadcast_to(v, (self.n,)), np.float32)
                for v in (rho, ux, uy, uz)]
        _check(lib().psl_diad_fill_equilibrium(self._h, *[_ptr(a) for a in arrs]))

    @property
    def esotwistFlipper(self) -> bool:
        v = ctypes.c_int()
        _check(lib().psl_diad_get_flipper(self._h, ctypes.byref(v)))
        return bool(v.value)

    def get_f(self) -> np.ndarray:
        out = np.empty((27, self.n), np.float32)
        _check(lib().psl_diad_get_f(self._h, out.ctypes.data))
        return out

    def set_f(self, f):
        a = np.ascontiguousarray(f, np.float32)
        _check(lib().psl_diad_set_f(self._h, a.ctypes.data))


# ------------------------------------------------------ DIAD grid builder
# psl_marker_rule (include/psl.h): the getMarkers callbacks of the drivers as
# an ordered rule list evaluated on the device during buildDIADGrids.
RULE_DTYPE = np.dtype([("what", "<i4"), ("value", "<i4"), ("level_mask", "<u4"), ("shape", "<i4"),
                       ("p", "<f4", (8,)), ("idx", "<i4", (6,))])
RULE_BOUNCEBACK, RULE_REFINEMENT, RULE_OUTLET = 0, 1, 2
SHAPE_SPHERE_LE, SHAPE_SPHERE_LT, SHAPE_XCYLINDER, SHAPE_BOX, SHAPE_INDEX_BOX = range(5)
DIAD_ARRAYS = ("ijk", "bounceback", "interface", "nbr", "c2f_read", "c2f_write", "f2c_write", "f2c_read", "outlet")
ALL_LEVELS = 0xFFFFFFFF


def marker_rule(what: int, shape: int, p=(), idx=(), value: int = 1, levels=ALL_LEVELS) -> np.ndarray:
    """One rule; `levels` is a bit mask of gridIDs or an iterable of them."""
    r = np.zeros(1, RULE_DTYPE)
    mask = levels if isinstance(levels, int) else sum(1 << int(l) for l in levels)
    r["what"], r["value"], r["level_mask"], r["shape"] = what, value, mask, shape
    r["p"][0, :len(p)] = p
    r["idx"][0, :len(idx)] = idx
    return r


def rules(*rs) -> np.ndarray:
    return np.concatenate(rs) if rs else np.zeros(0, RULE_DTYPE)


class DIADLevels:
    """buildDIADGrids (include/DIADFunctions.h:640-1046): the per-level cell
    tables produced on the device."""

    def __init__(self, base: InfoStruct, levels: int, wall_refinement_span: int, nu_phys: float,
                 stls: Sequence["STLStruct"] = (), marker_rules=None):
        require_device()
        r = np.zeros(0, RULE_DTYPE) if marker_rules is None else np.ascontiguousarray(marker_rules, RULE_DTYPE)
        arr = (ctypes.c_void_p * max(1, len(stls)))(*[s.handle.value for s in stls])
        ci = base.to_c()
        self.levels = levels
        self._h = ctypes.c_void_p()
        _check(lib().psl_diad_build(ctypes.byref(ci), levels, wall_refinement_span, nu_phys, arr, len(stls),
                                    r.ctypes.data if r.size else None, r.size, ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().psl_diad_levels_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def build_ms(self) -> float:
        v = ctypes.c_double()
        _check(lib().psl_diad_levels_build_ms(self._h, ctypes.byref(v)))
        return v.value

    def info(self, level: int):
        ci = _Info()
        counts = np.zeros(len(DIAD_ARRAYS), np.int64)
        _check(lib().psl_diad_levels_info(self._h, level, ctypes.byref(ci), counts.ctypes.data))
        return InfoStruct.from_c(ci), counts

    def array(self, level: int, name: str) -> np.ndarray:
        which = DIAD_ARRAYS.index(name)
        _, counts = self.info(level)
        a = np.zeros(int(counts[which]), np.uint8 if name in ("bounceback", "interface") else np.int32)
        if a.size:
            _check(lib().psl_diad_levels_get(self._h, level, which, a.ctypes.data))
        if name == "ijk":
            a = a.reshape(3, -1)
        elif name == "nbr":
            a = a.reshape(10, -1)
        return a

    def grids(self, storage: int | str = F32) -> list:
        return [DIADGridStruct.from_levels(self, l, storage) for l in range(self.levels)]

    def level(self, level: int) -> dict:
        info, _ = self.info(level)
        d = {"info": info, "n": info.cellCount}
        for name in DIAD_ARRAYS:
            d[name] = self.array(level, name)
        return d


def buildDIADGrids(base: InfoStruct, levels: int, wall_refinement_span: int, nu_phys: float,
                   stls: Sequence["STLStruct"] = (), marker_rules=None) -> DIADLevels:
    return DIADLevels(base, levels, wall_refinement_span, nu_phys, stls, marker_rules)


# ------------------------------------------------------- scalar transport
class ScalarTransportStruct:
    """include/types.h:121 on a periodic-shift GridStruct (FP32 storage)."""

    def __init__(self, grid: "GridStruct", tauT: float):
        self.grid = grid
        self.tauT = float(tauT)
        _check(lib().psl_scalar_enable(grid.handle, self.tauT))

    def set_given(self, given):
        a = None if given is None else np.ascontiguousarray(given, np.float32)
        _check(lib().psl_scalar_set_given(self.grid.handle, _ptr(a)))

    def get_T(self) -> np.ndarray:
        out = np.empty((27, self.grid.n), np.float32)
        _check(lib().psl_scalar_get(self.grid.handle, out.ctypes.data))
        return out

    def set_T(self, T):
        a = np.ascontiguousarray(T, np.float32)
        _check(lib().psl_scalar_set(self.grid.handle, a.ctypes.data))


def applyScalarTransportLocalCellUpdate(grid: "GridStruct", scalar: ScalarTransportStruct):
    """applyLocalCellUpdate(GridStruct&, ScalarTransportStruct&)."""
    _check(lib().psl_scalar_apply_local_cell_update(grid.handle))


def applyZeroGradientOutletScalarTransportBC(grid: "GridStruct", scalar: ScalarTransportStruct):
    """scalarTransport/scalarTransportMain.cu:95-135: the last x layer copies the
    scalar populations of the layer before it (between streaming and the update)."""
    _check(lib().psl_scalar_apply_zero_gradient_outlet_x(grid.handle))


# ------------------------------------------------------------ output
PLANES = {"XY": 0, "ZY": 1, "ZX": 2}


def section_cut(grid, plane: str, cut: int, rho_nominal: float = 1.225, soundspeed: float = 1.0,
                scalar_transport: bool = False, max_pixels: int = 20000000) -> np.ndarray:
    """exportSectionCutPlotGeneral on a GridStruct (rows x cols x 5; x 6 with
    scalar_transport, the last component the scalar) or a list of
    DIADGridStruct levels (rows x cols x 6, the last component the level)."""
    pl = PLANES[plane]
    if scalar_transport:
        dims = np.zeros(2, np.int32)
        _check(lib().psl_scalar_section_cut(grid.handle, pl, cut, rho_nominal, soundspeed, None, dims.ctypes.data))
        out = np.zeros((dims[0], dims[1], 6), np.float32)
        _check(lib().psl_scalar_section_cut(grid.handle, pl, cut, rho_nominal, soundspeed, out.ctypes.data,
                                            dims.ctypes.data))
        return out
    if isinstance(grid, (list, tuple)):
        arr = (ctypes.c_void_p * len(grid))(*[g.handle.value for g in grid])
        dims = np.zeros(3, np.int32)
        _check(lib().psl_diad_section_cut(arr, len(grid), pl, cut, rho_nominal, soundspeed, max_pixels, None,
                                          dims.ctypes.data))
        out = np.zeros((dims[0], dims[1], 6), np.float32)
        _check(lib().psl_diad_section_cut(arr, len(grid), pl, cut, rho_nominal, soundspeed, max_pixels,
                                          out.ctypes.data, dims.ctypes.data))
        return out
    dims = np.zeros(2, np.int32)
    _check(lib().psl_section_cut(grid.handle, pl, cut, rho_nominal, soundspeed, None, dims.ctypes.data))
    out = np.zeros((dims[0], dims[1], 5), np.float32)
    _check(lib().psl_section_cut(grid.handle, pl, cut, rho_nominal, soundspeed, out.ctypes.data, dims.ctypes.data))
    return out


def getAvgInletPressure(grid: "GridStruct", rho_nominal: float = 1.225, soundspeed: float = 1.0) -> float:
    """getAvgInletPressure (cascade/cascadeMain.cu:77-112): the pressure of the
    mean density over the j = 0 plane, every cell counted; the plane comes from
    the device section cut (pressure is linear in rho), averaged in FP64."""
    return float(section_cut(grid, "ZX", 0, rho_nominal, soundspeed)[..., 0].astype(np.float64).mean())


def getFlowReport(levels, plane: str, cut: int, bounds) -> dict:
    """getFlowReportXY/ZY/ZX over DIAD levels: mean ux, uy, uz, rho (lattice units)
    and the open area (mm^2) of the plane clipped to bounds (xmin, ymin, zmin,
    xmax, ymax, zmax)."""
    arr = (ctypes.c_void_p * len(levels))(*[g.handle.value for g in levels])
    b = np.ascontiguousarray(bounds, np.float32)
    out = np.zeros(5, np.float64)
    _check(lib().psl_diad_flow_report(arr, len(levels), PLANES[plane], cut, b.ctypes.data, out.ctypes.data))
    return {"ux": out[0], "uy": out[1], "uz": out[2], "rho": out[3], "areamm2": out[4]}


def exportSectionCutPlot(grid, plane: str, cut: int, plot_number: int, path: str = "/dev/shm/sim_data.bin",
                         rho_nominal: float = 1.225, soundspeed: float = 1.0,
                         scalar_transport: bool = False) -> np.ndarray:
    """exportSectionCutPlotXY/ZY/ZX: the cut written in the plotters' sim_data.bin
    format (six components with scalar_transport, as plotterScalarTransport.py reads)."""
    data = section_cut(grid, plane, cut, rho_nominal, soundspeed, scalar_transport)
    write_sim_data(path, plot_number, data)
    return data


def exportSectionCutPlotToiletPaperZ(grid: GridStruct, r: float, plot_number: int,
                                     path: Optional[str] = "/dev/shm/sim_data.bin", rho_nominal: float = 1.225,
                                     soundspeed: float = 1.0, max_pixels: int = 20000000) -> np.ndarray:
    """The cylinder of radius r (mm) about z unrolled: rows x cellCountZ x 5 of
    p, uz, u_tangential, u_radial (m/s) and the bounceback marker; over a list
    of DIAD levels rows x cols x 6, the last component the level."""
    if isinstance(grid, (list, tuple)):
        arr = (ctypes.c_void_p * len(grid))(*[g.handle.value for g in grid])
        dims = np.zeros(3, np.int32)
        _check(lib().psl_diad_cylinder_cut(arr, len(grid), r, rho_nominal, soundspeed, max_pixels, None,
                                           dims.ctypes.data))
        out = np.zeros((dims[0], dims[1], 6), np.float32)
        _check(lib().psl_diad_cylinder_cut(arr, len(grid), r, rho_nominal, soundspeed, max_pixels, out.ctypes.data,
                                           dims.ctypes.data))
        if path is not None:
            write_sim_data(path, plot_number, out)
        return out
    dims = np.zeros(2, np.int32)
    _check(lib().psl_cylinder_cut(grid.handle, r, rho_nominal, soundspeed, None, dims.ctypes.data))
    out = np.zeros((dims[0], dims[1], 5), np.float32)
    _check(lib().psl_cylinder_cut(grid.handle, r, rho_nominal, soundspeed, out.ctypes.data, dims.ctypes.data))
    if path is not None:
        write_sim_data(path, plot_number, out)
    return out


def export3DPlot(grid: GridStruct, plot_number: int, path: Optional[str] = "/dev/shm/sim_data.bin",
                 rho_nominal: float = 1.225, soundspeed: float = 1.0) -> np.ndarray:
    """export3DPlot: (nx, ny, nz, 5) of p, ux, uy, uz (m/s) and the bounceback
    marker, written to path (None: not written) in the 3-D sim_data.bin format."""
    I = grid.Info
    out = np.zeros((I.cellCountX, I.cellCountY, I.cellCountZ, 5), np.float32)
    _check(lib().psl_volume_export(grid.handle, rho_nominal, soundspeed, out.ctypes.data))
    if path is not None:
        _check(lib().psl_write_volume_data(path.encode(), plot_number, *out.shape, out.ctypes.data))
    return out


def write_sim_data(path: str, plot_number: int, data: np.ndarray):
    d = np.ascontiguousarray(data, np.float32)
    _check(lib().psl_write_sim_data(path.encode(), plot_number, d.shape[0], d.shape[1], d.shape[2], d.ctypes.data))


def exportHistoryData(values, current_iteration: int, path: str = "/dev/shm/historyData.bin"):
    """exportHistoryData (exportSectionCutPlot.h:994-1001): the first current_iteration + 1 values."""
    v = np.ascontiguousarray(values, np.float32)[: current_iteration + 1]
    _check(lib().psl_write_history(path.encode(), v.ctypes.data, v.size))


# ------------------------------------------------------------ operators
def fillEquilibriumFromFunction(grid: GridStruct, getInitialRhoUxUyUz: Optional[Callable] = None):
    """include/fillEquilibrium.h:1-22 with a vectorised host callback
    getInitialRhoUxUyUz(i, j, k, flags, Info) -> (rho, ux, uy, uz)."""
    if getInitialRhoUxUyUz is None:
        _check(lib().psl_fill_equilibrium(grid.handle, None, None, None, None))
        return
    i, j, k = grid.ijk()
    vals = getInitialRhoUxUyUz(i, j, k, grid.get_flags(), grid.Info)
    arrs = [None if v is None else np.ascontiguousarray(np.broadcast_to(v, (grid.n,)), np.float32) for v in vals]
    _check(lib().psl_fill_equilibrium(grid.handle, *[_ptr(a) for a in arrs]))


def fill_equilibrium(grid: GridStruct, rho=None, ux=None, uy=None, uz=None):
    arrs = [None if v is None else np.ascontiguousarray(np.broadcast_to(v, (grid.n,)), np.float32)
            for v in (rho, ux, uy, uz)]
    _check(lib().psl_fill_equilibrium(grid.handle, *[_ptr(a) for a in arrs]))


def synchronize():
    _check(lib().psl_device_synchronize())


class DeviceBuffer:
    """A raw device allocation (used for counter calibration runs)."""

    def __init__(self, nbytes: int):
        require_device()
        self.nbytes = nbytes
        self._p = ctypes.c_void_p()
        _check(lib().psl_device_alloc(nbytes, ctypes.byref(self._p)))

    @property
    def ptr(self):
        return self._p

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        buf = cls(max(1, a.nbytes))
        buf.upload(a)
        return buf

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        _check(lib().psl_device_upload(self._p, a.ctypes.data, a.nbytes))

    def download(self, shape, dtype=np.float32) -> np.ndarray:
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise ValueError("requested more bytes than the buffer holds")
        _check(lib().psl_device_download(out.ctypes.data, self._p, out.nbytes))
        return out

    def close(self):
        if getattr(self, "_p", None) and self._p.value:
            lib().psl_device_free(self._p)
            self._p = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


MLP_INPUTS, MLP_PARAMS, MLP_NU_SCALE = 9, 1409, 1e-3


class MLP:
    """Learned eddy-viscosity closure (psl_mlp_*): 9 -> 32 tanh -> 32 tanh -> 1."""

    def __init__(self, params=None, mean=None, std=None):
        require_device()
        self._h = ctypes.c_void_p()
        _check(lib().psl_mlp_create(ctypes.byref(self._h)))
        if params is not None:
            self.set_params(params, mean, std)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().psl_mlp_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params, mean=None, std=None):
        p = np.ascontiguousarray(params, np.float32)
        if p.size != MLP_PARAMS:
            raise ValueError(f"expected {MLP_PARAMS} parameters")
        m = None if mean is None else np.ascontiguousarray(mean, np.float32)
        s = None if std is None else np.ascontiguousarray(std, np.float32)
        _check(lib().psl_mlp_set_params(self._h, p.ctypes.data, None if m is None else m.ctypes.data,
                                        None if s is None else s.ctypes.data))

    def get_params(self):
        p = np.zeros(MLP_PARAMS, np.float32)
        m = np.zeros(MLP_INPUTS, np.float32)
        s = np.zeros(MLP_INPUTS, np.float32)
        _check(lib().psl_mlp_get_params(self._h, p.ctypes.data, m.ctypes.data, s.ctypes.data))
        return p, m, s

    def loss_grad(self, x_std: np.ndarray, t: np.ndarray):
        """FP32 training kernels: mean squared error and gradient on standardised inputs."""
        x = DeviceBuffer.from_array(np.asarray(x_std, np.float32))
        tt = DeviceBuffer.from_array(np.asarray(t, np.float32))
        loss = ctypes.c_double()
        g = np.zeros(MLP_PARAMS, np.float32)
        _check(lib().psl_mlp_loss_grad(self._h, x.ptr, tt.ptr, int(len(t)), ctypes.byref(loss), g.ctypes.data))
        return loss.value, g

    def train(self, raw_dev: DeviceBuffer, t_dev: DeviceBuffer, n: int, batch: int = 4096, iters: int = 500,
              lr: float = 1e-3, seed: int = 1):
        hist = np.zeros(iters, np.float32)
        _check(lib().psl_mlp_train(self._h, raw_dev.ptr, t_dev.ptr, int(n), int(batch), int(iters),
                                   ctypes.c_float(lr), ctypes.c_uint64(seed), hist.ctypes.data))
        return hist

    def eval(self, raw: np.ndarray) -> np.ndarray:
        """Network output of raw feature rows through the step's bf16 matrix-core path."""
        raw = np.ascontiguousarray(raw, np.float32).reshape(-1, MLP_INPUTS)
        d = DeviceBuffer.from_array(raw)
        y = DeviceBuffer(raw.shape[0] * 4)
        _check(lib().psl_mlp_eval(self._h, d.ptr, raw.shape[0], y.ptr))
        return y.download(raw.shape[0])


def les_filter(rho: DeviceBuffer, ux: DeviceBuffer, uy: DeviceBuffer, uz: DeviceBuffer, fine_nx: int, fine_ny: int,
               n: int, r: int) -> DeviceBuffer:
    """Box filter of the fine fields -> device buffer of 10 x m^3 coarse values."""
    m = n // r
    out = DeviceBuffer(10 * m ** 3 * 4)
    _check(lib().psl_les_filter(rho.ptr, ux.ptr, uy.ptr, uz.ptr, fine_nx, fine_ny, n, r, out.ptr))
    return out


def les_targets(coarse: DeviceBuffer, m: int, nu: float):
    """Raw features (m^3 x 9) and targets (nu_t / 1e-3) from filtered fields."""
    feat = DeviceBuffer(m ** 3 * MLP_INPUTS * 4)
    tgt = DeviceBuffer(m ** 3 * 4)
    _check(lib().psl_les_targets(coarse.ptr, m, ctypes.c_float(nu), feat.ptr, tgt.ptr))
    return feat, tgt


def probe_ps_pattern(nx: int, ny: int, nz: int, storage="f32", launches: int = 20, occupancy: int = 0) -> float:
    """HBM rate (GB/s, read + write) of the periodic-shift access pattern alone
    on this grid size, at `occupancy` workgroups per CU (0: no cap; the cap a
    grid's bulk kernel runs under, GridStruct.occupancy)."""
    if isinstance(storage, str):
        storage = STORAGE_NAMES[storage]
    require_device()
    g = ctypes.c_double()
    _check(lib().psl_probe_ps_pattern_capped(nx, ny, nz, storage, int(occupancy), launches, ctypes.byref(g)))
    return g.value


PROBE_MODES = {"inplace": 0, "read": 1, "write": 2}


def probe_ps_access(nx: int, ny: int, nz: int, storage="f32", streams: int = 27, mode: str = "inplace",
                    cells_per_lane: int = 1, line_aligned: bool = False, launches: int = 20) -> float:
    """HBM rate (GB/s of bytes moved) of the periodic-shift access pattern taken
    apart: `streams` populations, read + write in place / read only / write
    only, several consecutive cells per lane, shifts rounded to whole lines."""
    if isinstance(storage, str):
        storage = STORAGE_NAMES[storage]
    require_device()
    g = ctypes.c_double()
    _check(lib().psl_probe_ps_access(nx, ny, nz, storage, int(streams), PROBE_MODES[mode], int(cells_per_lane),
                                     int(bool(line_aligned)), launches, ctypes.byref(g)))
    return g.value


def device_copy(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int):
    _check(lib().psl_device_copy(dst.ptr, src.ptr, nbytes))


def applyStreaming(grid):
    """applyStreaming(GridStruct&) / applyStreaming(DIADGridStruct&)."""
    if isinstance(grid, DIADGridStruct):
        _check(lib().psl_diad_apply_streaming(grid.handle))
    else:
        _check(lib().psl_apply_streaming(grid.handle))


def applyLocalCellUpdate(grid):
    """applyLocalCellUpdate(GridStruct&) / applyLocalCellUpdate(DIADGridStruct&)."""
    if isinstance(grid, DIADGridStruct):
        _check(lib().psl_diad_apply_local_cell_update(grid.handle))
    else:
        _check(lib().psl_apply_local_cell_update(grid.handle))


def applyNonReflectiveOutletZ(grid: DIADGridStruct):
    """include/boundaryConditions/applyNonReflectiveOutlet.h:9-73 (before applyStreaming)."""
    _check(lib().psl_diad_apply_nonreflective_outlet_z(grid.handle))


def rotateSTLAlongX(stl: "STLStruct", radians: float):
    stl.rotate(0, radians)


def rotateSTLAlongY(stl: "STLStruct", radians: float):
    stl.rotate(1, radians)


def rotateSTLAlongZ(stl: "STLStruct", radians: float):
    stl.rotate(2, radians)


def getTorque(grid: DIADGridStruct, rho_nominal_phys: float) -> float:
    """getTorque (M35Full/m35Main.cu:176-260): torque about the z axis (N m) of
    the grid's forced-velocity cells."""
    return float(grid.forcing(rho_nominal_phys)[5])


def applyBCPeriodicX(grid: GridStruct):
    _check(lib().psl_apply_bc_periodic_x(grid.handle))


def applyBCPeriodicY(grid: GridStruct):
    _check(lib().psl_apply_bc_periodic_y(grid.handle))


def applyCoarseFineGridCommunication(coarse, fine):
    """GridStruct pair (include/gridRefinementFunctions.h:433-526) or DIADGridStruct pair (540-708)."""
    if isinstance(coarse, DIADGridStruct):
        _check(lib().psl_diad_coarse_fine_communication(coarse.handle, fine.handle))
    else:
        _check(lib().psl_coarse_fine_communication(coarse.handle, fine.handle))


def fillCoarseGridFromFine(coarse: GridStruct, fine: GridStruct):
    _check(lib().psl_fill_coarse_from_fine(coarse.handle, fine.handle))


def updateGrid(grids, steps: int = 1):
    """The level recursion over GridStruct or DIADGridStruct levels, `steps` coarse steps."""
    arr = (ctypes.c_void_p * len(grids))(*[g.handle.value for g in grids])
    if grids and isinstance(grids[0], DIADGridStruct):
        _check(lib().psl_diad_update_grid(arr, len(grids), steps))
    else:
        _check(lib().psl_update_grid(arr, len(grids), steps))


# ------------------------------------------------------------ geometry
class STLStruct:
    """include/types.h:84-119 (host and device copies)."""

    def __init__(self, handle: ctypes.c_void_p):
        self._h = handle

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().psl_stl_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    @property
    def triangleCount(self) -> int:
        c = ctypes.c_int()
        _check(lib().psl_stl_info(self._h, ctypes.byref(c), None))
        return c.value

    @property
    def bounds(self) -> np.ndarray:
        b = np.zeros(6, np.float32)
        _check(lib().psl_stl_info(self._h, None, b.ctypes.data))
        return b

    def set_origin(self, ox, oy, oz):
        _check(lib().psl_stl_set_origin(self._h, ox, oy, oz))

    def triangles(self) -> np.ndarray:
        out = np.zeros((self.triangleCount, 9), np.float32)
        _check(lib().psl_stl_get_triangles(self._h, out.ctypes.data))
        return out

    def rotate(self, axis: int, radians: float):
        """rotateSTLAlongX/Y/Z (axis 0, 1, 2) about the STL's own origin."""
        _check(lib().psl_stl_rotate(self._h, int(axis), radians))

    @staticmethod
    def from_triangles(tri) -> "STLStruct":
        t = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
        require_device()
        h = ctypes.c_void_p()
        _check(lib().psl_stl_from_triangles(t.ctypes.data, t.shape[0], ctypes.byref(h)))
        return STLStruct(h)


def readSTL(path: str) -> STLStruct:
    require_device()
    h = ctypes.c_void_p()
    _check(lib().psl_stl_read(path.encode(), ctypes.byref(h)))
    return STLStruct(h)


def checkSTLEdges(stl: STLStruct):
    bad = ctypes.c_int()
    counts = np.zeros((stl.triangleCount, 3), np.int32)
    _check(lib().psl_stl_check_edges(stl.handle, ctypes.byref(bad), counts.ctypes.data))
    return bad.value, counts


class BoolArray:
    """BoolArrayType on the device."""

    def __init__(self, n: int, value: bool = False):
        require_device()
        self.n = n
        self._h = ctypes.c_void_p()
        _check(lib().psl_mask_create(n, int(value), ctypes.byref(self._h)))

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            lib().psl_mask_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    def get(self) -> np.ndarray:
        out = np.empty(self.n, np.uint8)
        _check(lib().psl_mask_get(self._h, out.ctypes.data))
        return out

    def set(self, host):
        a = np.ascontiguousarray(host, np.uint8).reshape(-1)
        _check(lib().psl_mask_set(self._h, a.ctypes.data))


def applyMarkersInsideSTL(markerArray: BoolArray, stl: STLStruct, insideMarkerValue: bool, Info: InfoStruct):
    ci = Info.to_c()
    _check(lib().psl_markers_inside_stl(markerArray.handle, stl.handle, int(insideMarkerValue), ctypes.byref(ci)))


def multiplyBoolArrays(a: BoolArray, b: BoolArray, out: BoolArray):
    _check(lib().psl_mask_and(a.handle, b.handle, out.handle))


def sumBoolArrays(a: BoolArray, b: BoolArray, out: BoolArray):
    _check(lib().psl_mask_or(a.handle, b.handle, out.handle))


def invertBoolArray(a: BoolArray):
    _check(lib().psl_mask_not(a.handle))


def applyBouncebackMask(grid: GridStruct, mask: BoolArray):
    _check(lib().psl_grid_apply_bounceback_mask(grid.handle, mask.handle))
