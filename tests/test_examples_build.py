"""Every host driver under periodicshiftlbm_amd/examples compiles and links
against include/psl_facade.hpp and libpsl_amd.so with the host compiler alone
(the facade is plain C++ over the C ABI); the GPU tests run the built ones."""
import os
import tempfile

import pytest

from periodicshiftlbm_amd import build as B


@pytest.mark.parametrize("src", B.example_sources())
def test_example_compiles_and_links(src):
    if not os.path.exists(B.LIB):
        pytest.fail("libpsl_amd.so is not built (run __graft_entry__.build())")
    with tempfile.TemporaryDirectory() as d:
        exe = B.build_example(src, d)
        assert os.path.getsize(exe) > 0 and os.access(exe, os.X_OK)


def test_three_examples_present():
    assert B.example_sources() == ["flow_around_sphere.cpp", "flow_around_sphere_diad.cpp", "rotor_forcing_diad.cpp"]
