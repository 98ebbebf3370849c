"""The C-ABI library builds, loads and exports every symbol include/psl.h declares."""
import ctypes
import os
import re

REPO = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_functions():
    text = open(os.path.join(REPO, "include", "psl.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(psl_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import periodicshiftlbm_amd as P
    L = P.lib()
    names = declared_functions()
    assert len(names) > 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the Python mirror binds every declared entry point
    bound = {s[0] for s in P._SIGS}
    assert set(names) <= bound, set(names) - bound


def test_abi_version_and_error_path_without_gpu():
    import periodicshiftlbm_amd as P
    L = P.lib()
    assert L.psl_abi_version() == 1
    # a null grid is rejected with an error code and a message, no device needed
    rc = L.psl_apply_streaming(None)
    assert rc == -1
    assert b"null" in L.psl_last_error()


def test_oracle_library_is_test_only():
    """The product library does not link the oracle."""
    import periodicshiftlbm_amd as P
    path = P.LIB_PATH
    data = open(path, "rb").read()
    assert b"oracle_" not in data


def test_stl_reader_rejects_inconsistent_header_without_gpu(tmp_path):
    """A binary STL whose triangle count disagrees with its size (or exceeds
    the int range) is refused before anything is allocated for it."""
    import struct

    import periodicshiftlbm_amd as P
    L = P.lib()
    h = ctypes.c_void_p()
    for count, records in ((4_000_000_000, 1), (3, 2)):
        path = tmp_path / f"bad{count}.stl"
        path.write_bytes(b"\0" * 80 + struct.pack("<I", count) + b"\0" * (50 * records))
        rc = L.psl_stl_read(str(path).encode(), ctypes.byref(h))
        assert rc == -4, rc  # PSL_E_IO
        assert b"bytes on disk" in L.psl_last_error()


def test_size_and_argument_limits_without_gpu():
    """Sizes and arguments are checked before any device work, so the limits
    hold on a host without a GPU: empty and oversized grids (2^31 cells and
    more: the cell index is 32-bit), unknown storage, out-of-range DIAD
    neighbours, the batch layout and closure-interval ranges, and the restart
    entry points on a null grid."""
    import periodicshiftlbm_amd as P
    L = P.lib()
    h = ctypes.c_void_p()
    E_ARG = -1

    def create(nx, ny, nz, storage=P.F32):
        ci = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz).to_c()
        return L.psl_grid_create(ctypes.byref(ci), storage, ctypes.byref(h))

    assert create(0, 4, 4) == E_ARG and b"positive" in L.psl_last_error()
    assert create(4, -1, 4) == E_ARG
    assert create(2048, 1024, 1024) == E_ARG and b"2^31" in L.psl_last_error()  # exactly 2^31 cells
    assert create(4, 4, 4, storage=7) == E_ARG and b"storage" in L.psl_last_error()
    # a DIAD table whose neighbour leaves the grid is refused before any upload
    ci = P.InfoStruct(cellCountX=2, cellCountY=1, cellCountZ=1).to_c()
    ijk = (ctypes.c_int * 6)(0, 1, 0, 0, 0, 0)
    nbr = (ctypes.c_int * 14)(*([1, 0] * 6 + [2, 0]))
    assert L.psl_diad_create(ctypes.byref(ci), P.F32, 2, ijk, nbr, ctypes.byref(h)) == E_ARG
    assert b"neighbour" in L.psl_last_error()
    assert L.psl_grid_set_occupancy(None, 3) == E_ARG
    # layout, closure interval and restart entry points refuse null handles and bad ranges
    assert L.psl_batch_set_layout(None, 1, 0) == E_ARG
    assert L.psl_grid_set_closure_every(None, 2) == E_ARG
    n = ctypes.c_int64()
    assert L.psl_grid_state_bytes(None, ctypes.byref(n)) == E_ARG
    assert L.psl_grid_load_state(None, None, 0) == E_ARG
    assert L.psl_diad_load_state(None, None, 0) == E_ARG
    bad = ctypes.c_int64()
    assert L.psl_grid_check_health(None, ctypes.byref(bad), None, None) == E_ARG
    # the access probe checks its arguments before allocating
    gbs = ctypes.c_double()
    assert L.psl_probe_ps_access(64, 64, 64, P.F32, 5, 0, 1, 0, 1, ctypes.byref(gbs)) == E_ARG
    assert b"streams" in L.psl_last_error()
    # so does the capped pattern probe: caps outside 0 and 2-6 are refused
    for cap in (-1, 1, 7):
        assert L.psl_probe_ps_pattern_capped(64, 64, 64, P.F32, cap, 5, ctypes.byref(gbs)) == E_ARG
        assert b"workgroups per CU" in L.psl_last_error()
