"""CPU restatement (FP64, numpy) of the learned eddy-viscosity closure.

TEST INFRASTRUCTURE ONLY, like the rest of oracle/: tests/ use it to check
the HIP training, inference, filtering and target kernels.  The closure has
no counterpart in the reference; it replaces the Smagorinsky relaxation rate
of include/cellFunctions.h:157-198 (getOmegaLES) by a network output.

Pins (tests/test_closure_oracle.py):
  * mlp_grad: analytic gradients against central finite differences of
    mlp_loss in FP64 (every parameter);
  * features: rotation of the strain direction by a cube symmetry permutes
    the direction components and leaves magnitude, trace and the third
    invariant unchanged; a pure-trace tensor has zero deviatoric part;
  * les_filter: a constant field is unchanged and has zero subgrid stress;
    a field linear in one coordinate keeps zero subgrid stress only for the
    uniform component (closed form of the box variance of a linear ramp);
  * les_targets: with a subgrid stress of exactly -2 nu0 S (Smagorinsky form
    with a known eddy viscosity) the target is nu0 at every cell.

Network: 9 inputs -> 32 tanh -> 32 tanh -> 1 linear; output y is the eddy
viscosity in units of NU_SCALE (clipped at 0 when used).  Flat parameter
vector (1409 values): W1[9][32], b1[32], W2[32][32], b2[32], W3[32], b3.
Inputs are standardised with per-input mean / std stored with the network.
"""
from __future__ import annotations

import numpy as np

N_IN, N_HID = 9, 32
N_PARAM = N_IN * N_HID + N_HID + N_HID * N_HID + N_HID + N_HID + 1
NU_SCALE = 1e-3


# ------------------------------------------------------------ parameters
def unpack(p):
    p = np.asarray(p)
    o = 0
    W1 = p[o:o + N_IN * N_HID].reshape(N_IN, N_HID); o += N_IN * N_HID
    b1 = p[o:o + N_HID]; o += N_HID
    W2 = p[o:o + N_HID * N_HID].reshape(N_HID, N_HID); o += N_HID * N_HID
    b2 = p[o:o + N_HID]; o += N_HID
    W3 = p[o:o + N_HID]; o += N_HID
    b3 = p[o]
    return W1, b1, W2, b2, W3, b3


def init_params(seed: int = 0, dtype=np.float64):
    """Glorot-uniform weights, zero biases, from a fixed seed."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((N_IN, N_HID), (N_HID, N_HID)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        parts += [rng.uniform(-lim, lim, fan_in * fan_out), np.zeros(fan_out)]
    lim = np.sqrt(6.0 / (N_HID + 1))
    parts += [rng.uniform(-lim, lim, N_HID), np.zeros(1)]
    return np.concatenate(parts).astype(dtype)


# ------------------------------------------------------------ network
def mlp_forward(p, x):
    """x: (B, 9) standardised inputs -> (y (B,), cache)."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    h1 = np.tanh(x @ W1 + b1)
    h2 = np.tanh(h1 @ W2 + b2)
    y = h2 @ W3 + b3
    return y, (x, h1, h2)


def mlp_loss(p, x, t):
    y, _ = mlp_forward(p, x)
    return float(np.mean((y - t) ** 2))


def mlp_grad(p, x, t):
    """Gradient of mean((y - t)^2) with respect to the flat parameters."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    y, (x, h1, h2) = mlp_forward(p, x)
    B = x.shape[0]
    dy = 2.0 * (y - t) / B
    gW3 = h2.T @ dy
    gb3 = np.sum(dy)
    dz2 = np.outer(dy, W3) * (1.0 - h2 * h2)
    gW2 = h1.T @ dz2
    gb2 = dz2.sum(0)
    dz1 = (dz2 @ W2.T) * (1.0 - h1 * h1)
    gW1 = x.T @ dz1
    gb1 = dz1.sum(0)
    return np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3, [gb3]])


def adam_step(p, g, m, v, step, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update (Kingma & Ba 2015, algorithm 1); step counts from 1."""
    m = b1 * m + (1 - b1) * g
    v = b2 * v + (1 - b2) * g * g
    mh = m / (1 - b1 ** step)
    vh = v / (1 - b2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


# ------------------------------------------------------------ features
def raw_features(rho, pxx, pyy, pzz, pxy, pxz, pyz):
    """Nine closure inputs from the non-equilibrium second moments
    Pi = kappa - rho/3 I of a cell (kappa: central second moments):
      0  |dev(Pi)| / rho            (Frobenius norm, the Smagorinsky P / rho)
      1  tr(Pi) / rho
      2-7 dev(Pi) / |dev(Pi)| components xx, yy, zz, xy, xz, yz (0 if |dev| = 0)
      8  3 sqrt(6) det(dev(Pi)) / |dev(Pi)|^3 in [-1, 1] (0 if |dev| = 0)."""
    rho = np.asarray(rho, np.float64)
    p = [np.asarray(a, np.float64) / rho for a in (pxx, pyy, pzz, pxy, pxz, pyz)]
    tr = p[0] + p[1] + p[2]
    dxx, dyy, dzz = p[0] - tr / 3, p[1] - tr / 3, p[2] - tr / 3
    dxy, dxz, dyz = p[3], p[4], p[5]
    mag = np.sqrt(dxx * dxx + dyy * dyy + dzz * dzz + 2 * (dxy * dxy + dxz * dxz + dyz * dyz))
    inv = np.where(mag > 0, 1.0 / np.where(mag > 0, mag, 1.0), 0.0)
    det = (dxx * (dyy * dzz - dyz * dyz) - dxy * (dxy * dzz - dyz * dxz) + dxz * (dxy * dyz - dyy * dxz))
    r = 3 * np.sqrt(6.0) * det * inv ** 3
    return np.stack([mag, tr, dxx * inv, dyy * inv, dzz * inv, dxy * inv, dxz * inv, dyz * inv, r], axis=-1)


def standardise(feat, mean, std):
    return (feat - mean) / std


# ------------------------------------------------------------ data
def les_filter(rho, ux, uy, uz, r: int):
    """Box filter of width r over a periodic n^3 field (arrays indexed
    [k, j, i]): returns filtered rho, Favre velocity and the subgrid stress
    tau_ab = <rho u_a u_b> / <rho> - u~_a u~_b (order xx, yy, zz, xy, xz, yz)."""
    def box(a):
        n = a.shape[0]
        m = n // r
        return a.reshape(m, r, m, r, m, r).mean(axis=(1, 3, 5))
    rb = box(rho)
    u = [box(rho * c) / rb for c in (ux, uy, uz)]
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    comps = (ux, uy, uz)
    tau = [box(rho * comps[a] * comps[b]) / rb - u[a] * u[b] for a, b in pairs]
    return rb, u[0], u[1], u[2], np.stack(tau)


def strain_rate(ux, uy, uz):
    """S_ab = (d_a u_b + d_b u_a) / 2 by central differences on a periodic grid
    (arrays [k, j, i]: axis 2 = x, 1 = y, 0 = z); order xx, yy, zz, xy, xz, yz."""
    def d(f, axis):
        return 0.5 * (np.roll(f, -1, axis) - np.roll(f, 1, axis))
    ax = {0: 2, 1: 1, 2: 0}  # x, y, z -> array axis
    u = (ux, uy, uz)
    g = [[d(u[b], ax[a]) for b in range(3)] for a in range(3)]  # g[a][b] = d_a u_b
    pairs = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))
    return np.stack([0.5 * (g[a][b] + g[b][a]) for a, b in pairs])


def les_targets(rho, tau, S, nu):
    """Eddy-viscosity target and the closure inputs a perfect closure would see.

    nu_t* = max(0, -dev(tau):S / (2 S:S)) (least-squares fit of dev(tau) = -2 nu_t S);
    Pi = -(2/3) rho tau_eff S with tau_eff = 3 (nu + nu_t*) + 1/2 (the
    Chapman-Enskog non-equilibrium second moment in lattice units)."""
    tr = (tau[0] + tau[1] + tau[2]) / 3
    dev = np.stack([tau[0] - tr, tau[1] - tr, tau[2] - tr, tau[3], tau[4], tau[5]])
    w = np.array([1, 1, 1, 2, 2, 2], np.float64).reshape(6, *([1] * (S.ndim - 1)))
    num = np.sum(w * dev * S, axis=0)
    den = 2 * np.sum(w * S * S, axis=0)
    nut = np.where(den > 0, np.maximum(0.0, -num / np.where(den > 0, den, 1.0)), 0.0)
    tau_eff = 3 * (nu + nut) + 0.5
    Pi = -(2.0 / 3.0) * rho * tau_eff * S
    return nut, raw_features(rho, *Pi)


def energy_spectrum(ux, uy, uz):
    """Shell-summed kinetic energy spectrum E(k), k = 0..n/2 (periodic n^3)."""
    n = ux.shape[0]
    e = np.zeros(n // 2 + 1)
    k = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(k[:, None, None] ** 2 + k[None, :, None] ** 2 + k[None, None, :] ** 2)
    shell = np.rint(kk).astype(int)
    for c in (ux, uy, uz):
        F = np.fft.fftn(c) / c.size
        np.add.at(e, np.minimum(shell, n // 2).ravel(), 0.5 * (np.abs(F) ** 2).ravel())
    return e


def spectrum_error(e, e_ref, kmax: int):
    """Mean |log(E / E_ref)| over shells 1..kmax."""
    k = np.arange(1, kmax + 1)
    return float(np.mean(np.abs(np.log(e[k] / e_ref[k]))))
