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
