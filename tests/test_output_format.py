"""Plotter file formats (include/plotter/exportSectionCutPlot.h:64-86, 994-1001).

The library's writers are host code; the files are read back the way the
reference's plotters read them (include/plotter/plotter.py:8-14,
flowAroundSphereDIAD/historyPlotter.py:44-47): an int32 plot number, three
int32 dimensions (rows, cols, components), then float32 rows; and an int32
count followed by count float32 values.
"""
import numpy as np

import periodicshiftlbm_amd as P


def test_sim_data_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    data = rng.standard_normal((7, 11, 6)).astype(np.float32)
    path = str(tmp_path / "sim_data.bin")
    P.write_sim_data(path, 42, data)
    with open(path, "rb") as f:
        plot_number = np.fromfile(f, dtype=np.int32, count=1)[0]
        n_vertical, n_horizontal, n_vars = np.fromfile(f, dtype=np.int32, count=3)
        back = np.fromfile(f, dtype=np.float32).reshape((n_vertical, n_horizontal, n_vars))
    assert plot_number == 42 and (n_vertical, n_horizontal, n_vars) == (7, 11, 6)
    assert np.array_equal(back, data)
    assert (tmp_path / "sim_data.bin").stat().st_size == 16 + data.nbytes


def test_history_round_trip(tmp_path):
    values = np.linspace(0.3, 0.5, 50, dtype=np.float32)
    path = str(tmp_path / "historyData.bin")
    P.exportHistoryData(values, 19, path)  # the first 20 values, as at iteration 19
    with open(path, "rb") as f:
        count = np.fromfile(f, dtype=np.int32, count=1)[0]
        back = np.fromfile(f, dtype=np.float32, count=count)
    assert count == 20 and np.array_equal(back, values[:20])


def test_writer_reports_bad_path():
    try:
        P.write_sim_data("/nonexistent-dir/x.bin", 0, np.zeros((1, 1, 5), np.float32))
    except P.PslError as e:
        assert "cannot open" in str(e)
    else:
        raise AssertionError("no error for an unwritable path")


def test_volume_data_round_trip(tmp_path):
    """export3DPlot's file (exportSectionCutPlot.h:640-670): int32 plot number,
    nx, ny, nz, components, then the cells in (i, j, k) order, k fastest."""
    rng = np.random.default_rng(1)
    data = rng.standard_normal((5, 4, 3, 5)).astype(np.float32)
    path = str(tmp_path / "sim_data.bin")
    P._check(P.lib().psl_write_volume_data(path.encode(), 9, 5, 4, 3, 5, data.ctypes.data))
    with open(path, "rb") as f:
        hdr = np.fromfile(f, dtype=np.int32, count=5)
        back = np.fromfile(f, dtype=np.float32)
    assert list(hdr) == [9, 5, 4, 3, 5] and back.size == data.size
    # the reference's writer loops i, then j, then k innermost, 5 floats per cell
    assert np.array_equal(back.reshape(5, 4, 3, 5)[2, 1, 0], data[2, 1, 0])
    assert np.array_equal(back, data.reshape(-1))
