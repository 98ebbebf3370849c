"""Scalar transport on the device (psl_scalar_*) against the oracle.

A sphere channel (48 x 24 x 24: velocity inlet, pressure outlet, bounceback
sphere, Smagorinsky 0.1) carries a scalar prescribed at the inlet; ten steps
of streaming + applyLocalCellUpdate(Grid, ScalarTransport) on the GPU and in
the oracle give f and T within 1e-5 relative (FP32).
"""
import numpy as np
import pytest

import oracle as O
from periodicshiftlbm_amd import cases

pytestmark = pytest.mark.gpu


def _rel_err(a, b):
    return float(np.max(np.abs(a.astype(np.float64) - b.astype(np.float64)) / np.maximum(np.abs(b), 1e-3)))


def test_scalar_transport_matches_oracle(psl):
    P = psl
    shape = nx, ny, nz = (48, 24, 24)
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = ((i3 - 12.0) ** 2 + (j3 - 11.5) ** 2 + (k3 - 11.5) ** 2 <= 16.0).reshape(-1)
    flags = cases.channel_flags(shape, solid)
    bc = cases.channel_given_values(0.04)
    nu, cs, tau_t = 0.02, 0.1, 0.7
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=nu)
    g = P.GridStruct(info, "f32")
    g.set_flags(flags)
    g.set_given_values(bc)
    g.set_smagorinsky(cs)
    P.fill_equilibrium(g, ux=np.full(g.n, 0.04, np.float32))
    st = P.ScalarTransportStruct(g, tau_t)
    given = np.full(g.n, np.nan, np.float32)
    given[i3.reshape(-1) == 0] = 1.0
    st.set_given(given)
    og = O.Grid(nx, ny, nz, dtype=np.float32, nu=nu, cs=cs, flags=flags, bcvals=bc)
    og.f[:] = g.get_f()
    T = np.zeros((27, g.n), np.float32)
    for step in range(10):
        P.applyStreaming(g)
        P.applyScalarTransportLocalCellUpdate(g, st)
        og.stream()
        og.scalar_update(T, tau_t, given)
    n = g.n
    Tl = np.stack([T[q][(np.arange(n) + og.shift[q]) % n] for q in range(27)])
    ef, eT = _rel_err(g.get_f(), og.logical_f()), _rel_err(st.get_T(), Tl)
    conc = (og.logical_f() * Tl).sum(0) / og.logical_f().sum(0)
    print(f"[scalar] f rel err {ef:.3e}, T rel err {eT:.3e}, scalar reached i = "
          f"{int(np.max(np.nonzero(conc.reshape(nz, ny, nx).max(axis=(0, 1)) > 1e-3)))}", flush=True)
    assert ef < 1e-5 and eT < 1e-5
    assert g.get_f().dtype == np.float32


def test_scalar_transport_with_zero_gradient_outlet(psl):
    """The scalar driver's step (streaming, zero-gradient scalar outlet at the
    last x layer, update) against the oracle's, 12 steps, f and T within 1e-5."""
    P = psl
    shape = nx, ny, nz = (40, 20, 20)
    k3, j3, i3 = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    solid = ((i3 - 10.0) ** 2 + (j3 - 9.5) ** 2 + (k3 - 9.5) ** 2 <= 9.0).reshape(-1)
    flags = cases.channel_flags(shape, solid)
    bc = cases.channel_given_values(0.05)
    nu, cs, tau_t = 0.02, 0.1, 0.6
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=nu)
    g = P.GridStruct(info, "f32")
    g.set_flags(flags)
    g.set_given_values(bc)
    g.set_smagorinsky(cs)
    P.fill_equilibrium(g, ux=np.full(g.n, 0.05, np.float32))
    st = P.ScalarTransportStruct(g, tau_t)
    given = np.full(g.n, np.nan, np.float32)
    given[i3.reshape(-1) == 0] = 1.0
    st.set_given(given)
    og = O.Grid(nx, ny, nz, dtype=np.float32, nu=nu, cs=cs, flags=flags, bcvals=bc)
    og.f[:] = g.get_f()
    # a scalar already spread over the box, so the outlet layer carries values
    # (both grids are unstreamed: raw and logical order agree)
    T = (0.5 + 0.4 * np.sin(0.3 * i3 + 0.2 * j3).reshape(1, -1) *
         np.ones((27, 1))).astype(np.float32)
    st.set_T(T)
    for step in range(12):
        P.applyStreaming(g)
        P.applyZeroGradientOutletScalarTransportBC(g, st)
        P.applyScalarTransportLocalCellUpdate(g, st)
        og.stream()
        og.scalar_outlet_x(T)
        og.scalar_update(T, tau_t, given)
    n = g.n
    Tl = np.stack([T[q][(np.arange(n) + og.shift[q]) % n] for q in range(27)])
    ef, eT = _rel_err(g.get_f(), og.logical_f()), _rel_err(st.get_T(), Tl)
    print(f"[scalar] zero-gradient outlet: f rel err {ef:.3e}, T rel err {eT:.3e}", flush=True)
    assert ef < 1e-5 and eT < 1e-5
