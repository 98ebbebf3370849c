"""Scalar transport in the oracle (include/scalarTransportFunctions.h, Zhang, Fan,
Chen 2011): CPU pins of applyScalarTransportCollision and of the update.

- a uniform scalar is a fixed point;
- the collision keeps the scalar content sum_q f_q T_q;
- with f an equilibrium the scalar flux deviation sum_q c_q f_q (T_q - Tsum)
  relaxes by exactly (1 - 1/tauT);
- on a box at rest a sinusoidal scalar decays at the diffusivity (tauT - 1/2)/3.
FP64 instantiation.
"""
import numpy as np
import pytest

import oracle as O


def _feq(n, rng, u_scale=0.05):
    ru = np.column_stack([1 + 0.05 * rng.standard_normal(n), u_scale * rng.standard_normal((n, 3))])
    return O.feq(ru)


def test_uniform_scalar_is_fixed_point():
    rng = np.random.default_rng(0)
    f = _feq(50, rng) + 1e-3 * rng.standard_normal((50, 27))
    T = np.full((50, 27), 0.7)
    assert np.max(np.abs(O.scalar_collide(f, T, 0.83) - 0.7)) < 1e-14


@pytest.mark.parametrize("tau", [0.55, 0.8, 1.7])
def test_scalar_content_and_flux_relaxation(tau):
    rng = np.random.default_rng(1)
    c = np.array(O.velocity_set(), np.float64)
    f = _feq(200, rng)
    T = rng.random((200, 27))
    T2 = O.scalar_collide(f, T, tau)
    assert np.max(np.abs((f * T2).sum(1) - (f * T).sum(1))) < 1e-13
    tsum = (f * T).sum(1) / f.sum(1)
    flux_in = (f * (T - tsum[:, None])) @ c.T
    flux_out = (f * (T2 - tsum[:, None])) @ c.T
    assert np.max(np.abs(flux_out - (1 - 1 / tau) * flux_in)) < 1e-13
    # a non-equilibrium f keeps the content too
    g = f + 1e-3 * rng.standard_normal(f.shape)
    assert np.max(np.abs((g * O.scalar_collide(g, T, tau)).sum(1) - (g * T).sum(1))) < 1e-13


def test_sinusoidal_scalar_decays_at_the_diffusivity():
    """The field varies along x with the period of the box: the shifted layout
    wraps x rows into the next row (and y into the next plane), which a field
    periodic in x and uniform in y, z does not see."""
    nx, ny, nz, tau, steps = 32, 4, 4, 0.8, 200
    g = O.Grid(nx, ny, nz, dtype=np.float64, nu=0.05, cs=0.0)
    g.fill_equilibrium()
    i = np.arange(nx * ny * nz) % nx
    amp0 = 0.1
    T = np.tile(1.0 + amp0 * np.sin(2 * np.pi * i / nx), (27, 1))  # shift 0: raw = logical
    amps = []
    for s in range(steps):
        g.stream()
        g.scalar_update(T, tau)
        if s in (steps // 2 - 1, steps - 1):
            f = g.logical_f()
            n = f.shape[1]
            Tl = np.stack([T[q][(np.arange(n) + g.shift[q]) % n] for q in range(27)])
            conc = ((f * Tl).sum(0) / f.sum(0)).reshape(-1, nx).mean(0)
            amps.append(2 * np.abs(np.fft.rfft(conc)[1]) / nx)
    d = np.log(amps[0] / amps[1]) / (steps // 2) / (2 * np.pi / nx) ** 2
    print(f"diffusivity {d:.5f} vs (tau - 1/2) / 3 = {(tau - 0.5) / 3:.5f}")
    assert abs(d / ((tau - 0.5) / 3) - 1) < 0.01


def test_fp32_instantiation_agrees_with_fp64():
    rng = np.random.default_rng(3)
    f = _feq(100, rng) + 1e-3 * rng.standard_normal((100, 27))
    T = rng.random((100, 27))
    a = O.scalar_collide(f, T, 0.7)
    b = O.scalar_collide(f.astype(np.float32), T.astype(np.float32), 0.7)
    assert np.max(np.abs(a - b)) < 1e-5


def test_zero_gradient_outlet_copies_the_layer_before():
    """applyZeroGradientOutletScalarTransportBC in the oracle on a streamed grid
    (non-zero shifts): in logical order, every non-bounceback cell of the last x
    layer holds the populations of its -x neighbour afterwards; bounceback cells
    and every other layer keep theirs."""
    nx, ny, nz = 7, 5, 4
    rng = np.random.default_rng(9)
    flags = np.zeros(nx * ny * nz, np.uint8)
    flags[rng.choice(nx * ny * nz, 30, replace=False)] = 1  # bounceback
    g = O.Grid(nx, ny, nz, dtype=np.float64, flags=flags)
    for _ in range(3):
        g.stream()
    assert np.any(g.shift != 0)
    n = g.n
    T = rng.random((27, n))

    def logical(a):
        return np.stack([a[q][(np.arange(n) + g.shift[q]) % n] for q in range(27)])

    before = logical(T).reshape(27, nz, ny, nx)
    g.scalar_outlet_x(T)
    after = logical(T).reshape(27, nz, ny, nx)
    bb = (flags.reshape(nz, ny, nx) & 7) == 1
    last = ~bb[:, :, -1]
    assert np.array_equal(after[:, :, :, -1][:, last], before[:, :, :, -2][:, last])
    assert np.array_equal(after[:, :, :, -1][:, ~last], before[:, :, :, -1][:, ~last])
    assert np.array_equal(after[:, :, :, :-1], before[:, :, :, :-1])
