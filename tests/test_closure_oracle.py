"""Pins of the CPU closure oracle (oracle/closure.py) on the CPU."""
import itertools

import numpy as np

from oracle import closure as C


def test_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = C.init_params(5) + rng.normal(0, 0.05, C.N_PARAM)
    x = rng.normal(0, 1, (17, C.N_IN))
    t = rng.normal(0, 1, 17)
    g = C.mlp_grad(p, x, t)
    h = 1e-6
    fd = np.empty(C.N_PARAM)
    for i in range(C.N_PARAM):
        e = np.zeros(C.N_PARAM)
        e[i] = h
        fd[i] = (C.mlp_loss(p + e, x, t) - C.mlp_loss(p - e, x, t)) / (2 * h)
    err = np.max(np.abs(g - fd)) / np.max(np.abs(fd))
    assert err < 1e-7, err


def test_adam_moves_against_gradient_and_fits():
    rng = np.random.default_rng(4)
    x = rng.normal(0, 1, (256, C.N_IN))
    t = np.tanh(x[:, 0]) + 0.5 * x[:, 1] ** 2
    p = C.init_params(1)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0 = C.mlp_loss(p, x, t)
    for s in range(1, 301):
        p, m, v = C.adam_step(p, C.mlp_grad(p, x, t), m, v, s, 1e-2)
    assert C.mlp_loss(p, x, t) < 0.1 * l0


def _rand_sym(rng, n):
    a = rng.normal(0, 1e-3, (6, n))
    return a


def test_features_invariants_under_cube_symmetry():
    rng = np.random.default_rng(5)
    P = _rand_sym(rng, 50)
    rho = rng.uniform(0.9, 1.1, 50)
    f0 = C.raw_features(rho, *P)
    # full tensor, permuted axes
    idx = {(0, 0): 0, (1, 1): 1, (2, 2): 2, (0, 1): 3, (1, 0): 3, (0, 2): 4, (2, 0): 4, (1, 2): 5, (2, 1): 5}
    for perm in itertools.permutations(range(3)):
        Q = np.stack([P[idx[(perm[a], perm[b])]] for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))])
        f1 = C.raw_features(rho, *Q)
        np.testing.assert_allclose(f1[:, [0, 1, 8]], f0[:, [0, 1, 8]], rtol=1e-10, atol=1e-14)
        # direction components are permuted the same way
        Qd = np.stack([f0[:, 2 + idx[(perm[a], perm[b])]] for a, b in
                       ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))], axis=-1)
        np.testing.assert_allclose(f1[:, 2:8], Qd, rtol=1e-10, atol=1e-14)
    # |r| <= 1 and the unit direction has Frobenius norm 1
    assert np.all(np.abs(f0[:, 8]) <= 1 + 1e-12)
    d = f0[:, 2:8]
    norm = d[:, 0] ** 2 + d[:, 1] ** 2 + d[:, 2] ** 2 + 2 * (d[:, 3] ** 2 + d[:, 4] ** 2 + d[:, 5] ** 2)
    np.testing.assert_allclose(norm, 1.0, rtol=1e-12)


def test_features_pure_trace():
    f = C.raw_features(np.ones(3), *np.array([[1e-3] * 3, [1e-3] * 3, [1e-3] * 3, [0] * 3, [0] * 3, [0] * 3]))
    np.testing.assert_allclose(f[:, 0], 0, atol=1e-18)
    np.testing.assert_allclose(f[:, 1], 3e-3)
    np.testing.assert_allclose(f[:, 2:], 0, atol=1e-18)


def test_filter_constant_and_ramp():
    n, r = 8, 4
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    rho = np.full((n, n, n), 1.2)
    ux = np.full((n, n, n), 0.03)
    rb, a, b, c, tau = C.les_filter(rho, ux, 0 * ux, 0 * ux, r)
    np.testing.assert_allclose(rb, 1.2)
    np.testing.assert_allclose(a, 0.03)
    np.testing.assert_allclose(tau, 0, atol=1e-18)
    # u_x = g * i: box variance of a ramp over r cells is g^2 (r^2 - 1) / 12
    g = 0.01
    rb, a, b, c, tau = C.les_filter(np.ones((n, n, n)), g * i.astype(float), 0 * ux, 0 * ux, r)
    np.testing.assert_allclose(tau[0], g * g * (r * r - 1) / 12, rtol=1e-12)
    np.testing.assert_allclose(tau[1:], 0, atol=1e-18)


def test_targets_recover_known_eddy_viscosity():
    rng = np.random.default_rng(6)
    n = 6
    S = rng.normal(0, 1e-3, (6, n, n, n))
    tr = (S[0] + S[1] + S[2]) / 3
    S[0] -= tr; S[1] -= tr; S[2] -= tr  # incompressible strain
    nu0 = 2.5e-3
    tau = -2 * nu0 * S
    tau[:3] += 1e-4  # an isotropic part does not change the target
    rho = np.ones((n, n, n))
    nut, feat = C.les_targets(rho, tau, S, nu=1e-4)
    np.testing.assert_allclose(nut, nu0, rtol=1e-10)
    assert feat.shape == (n, n, n, C.N_IN)


def test_spectrum_of_single_mode():
    n = 16
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    u = 0.1 * np.sin(2 * np.pi * 3 * i / n)
    e = C.energy_spectrum(u, 0 * u, 0 * u)
    # E(3) = 1/2 <u^2> = 0.0025, nothing elsewhere
    np.testing.assert_allclose(e[3], 0.5 * np.mean(u * u), rtol=1e-12)
    assert np.sum(e) - e[3] < 1e-20
