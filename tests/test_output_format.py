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
