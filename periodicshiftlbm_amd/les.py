"""Learned-closure workflow: decaying isotropic turbulence on a fine periodic
grid, filtered training data, training, and coarse LES runs judged against
the filtered fine run.

No reference counterpart (the reference's LES model is Smagorinsky,
include/cellFunctions.h:157-198).  All solver work runs through the library:
the fine and coarse steps, the filter / target kernels (psl_les_filter,
psl_les_targets), training (psl_mlp_train) and the step with the network
(psl_grid_set_closure).  The host does the initial random field (FFT) and
the energy spectra of the coarse velocity fields.

Scaling between the grids (same lattice velocity, acoustic scaling): a coarse
cell is r fine cells, a coarse step is r fine steps and nu_coarse = nu_fine / r.
"""
from __future__ import annotations

import os
import time

import numpy as np

from . import cases


def isotropic_field(n: int, k_peak: float, u_rms: float = 0.04, seed: int = 7):
    """Random-phase solenoidal velocity on a periodic n^3 box with shell
    spectrum E(k) ~ k^4 exp(-2 (k / k_peak)^2), scaled to u_rms (per
    component).  Arrays indexed [k, j, i] (z, y, x), float32."""
    import scipy.fft as sfft

    rng = np.random.default_rng(seed)
    k1 = np.fft.fftfreq(n, 1.0 / n).astype(np.float32)
    kz, ky, kx = np.meshgrid(k1, k1, k1, indexing="ij")
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    kk[0, 0, 0] = 1.0
    amp = np.sqrt(kk ** 4 * np.exp(-2.0 * (kk / k_peak) ** 2) / (4 * np.pi * kk * kk)).astype(np.float32)
    amp[0, 0, 0] = 0.0
    comps = []
    for _ in range(3):
        c = (rng.standard_normal((n, n, n), dtype=np.float32)
             + 1j * rng.standard_normal((n, n, n), dtype=np.float32)).astype(np.complex64)
        comps.append(c * amp)
    # divergence-free projection u - k (k.u) / k^2
    dot = (kx * comps[0] + ky * comps[1] + kz * comps[2]) / (kk * kk)
    comps = [comps[0] - kx * dot, comps[1] - ky * dot, comps[2] - kz * dot]
    del dot
    workers = max(1, min(16, os.cpu_count() or 1))
    out = []
    for c in comps:
        out.append(np.ascontiguousarray(sfft.ifftn(c, workers=workers).real.astype(np.float32)))
    rms = np.sqrt(np.mean([np.mean(u.astype(np.float64) ** 2) for u in out]))
    return [u * np.float32(u_rms / rms) for u in out]


def glorot_params(seed: int) -> np.ndarray:
    """Initial flat parameters (W1, b1, W2, b2, W3, b3): Glorot-uniform weights,
    zero biases, from a fixed seed."""
    rng = np.random.default_rng(seed)
    parts = []
    for fan_in, fan_out in ((9, 32), (32, 32), (32, 1)):
        lim = np.sqrt(6.0 / (fan_in + fan_out))
        parts += [rng.uniform(-lim, lim, fan_in * fan_out), np.zeros(fan_out)]
    return np.concatenate(parts).astype(np.float32)


def _padded(a: np.ndarray) -> np.ndarray:
    """n^3 periodic field -> the (n, n+1, n+1) array of the duplicated-layer layout."""
    return np.concatenate([np.concatenate([a, a[:, :1, :]], axis=1),
                           np.concatenate([a, a[:, :1, :]], axis=1)[:, :, :1]], axis=2)


def periodic_grid(P, n: int, nu: float, storage: str, u, cs: float = 0.0, rho=None):
    """Fully periodic n^3 box (duplicated x / y layers, z through the shift) at feq(rho, u)."""
    nx, ny, nz = cases.tgv_shape(n)
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=nu)
    g = P.GridStruct(info, storage)
    g.set_periodic_copies(x=True, y=True)
    if cs:
        g.set_smagorinsky(cs)
    r = np.ones((n, n, n), np.float32) if rho is None else rho
    P.fill_equilibrium(g, *[_padded(np.asarray(a, np.float32)).ravel() for a in (r, *u)])
    return g


def unique_velocity(g, n: int):
    r, a, b, c = g.macroscopic()
    sh = (n, n + 1, n + 1)
    return [x.reshape(sh)[:, :n, :n].astype(np.float64) for x in (a, b, c)]


def box_transfer2(m: int, r: int):
    """|G(k)|^2 on an m^3 coarse grid of the r-cell box average of the fine
    grid: per axis |sin(pi k / m) / (r sin(pi k / (r m)))|^2."""
    k1 = np.abs(np.fft.fftfreq(m, 1.0 / m))
    g = np.ones(m)
    nz = k1 > 0
    g[nz] = (np.sin(np.pi * k1[nz] / m) / (r * np.sin(np.pi * k1[nz] / (r * m)))) ** 2
    return g[:, None, None] * g[None, :, None] * g[None, None, :]


def energy_spectrum(ux, uy, uz, weight=None):
    """Shell-summed kinetic energy spectrum of a periodic n^3 field, k = 0..n/2;
    each mode's energy is multiplied by `weight` when given."""
    n = ux.shape[0]
    k1 = np.fft.fftfreq(n, 1.0 / n)
    kk = np.sqrt(k1[:, None, None] ** 2 + k1[None, :, None] ** 2 + k1[None, None, :] ** 2)
    shell = np.minimum(np.rint(kk).astype(int), n // 2).ravel()
    e = np.zeros(n // 2 + 1)
    for c in (ux, uy, uz):
        F = np.abs(np.fft.fftn(c) / c.size) ** 2
        if weight is not None:
            F = F * weight
        np.add.at(e, shell, 0.5 * F.ravel())
    return e


def spectrum_error(e, e_ref, kmax: int) -> float:
    """Mean |ln(E / E_ref)| over shells 1..kmax."""
    k = np.arange(1, kmax + 1)
    return float(np.mean(np.abs(np.log(e[k] / e_ref[k]))))


class Workflow:
    """Fine run -> filtered snapshots (training pairs + reference spectra) ->
    training -> coarse LES comparisons."""

    def __init__(self, P, n: int = 512, r: int = 4, k_peak: float = 8.0, u_rms: float = 0.04,
                 nu_fine: float = 5e-4, cs_fine: float = 0.1, storage: str = "f16", seed: int = 7, log=print,
                 closure_every: int = 1):
        self.P, self.n, self.r, self.m = P, n, r, n // r
        self.k_peak, self.u_rms, self.nu_fine, self.cs_fine = k_peak, u_rms, nu_fine, cs_fine
        self.nu_coarse = nu_fine / r
        self.storage, self.seed, self.log = storage, seed, log
        self.closure_every = closure_every  # network evaluated every this many coarse steps
        self.snapshots = {}   # fine step -> (coarse rho, u) host arrays of the filtered field
        self.feats, self.tgts = [], []  # device buffers per snapshot
        self.timings = {}

    def run_fine(self, steps: int, every: int, first_sample: int):
        P, n, r, m = self.P, self.n, self.r, self.m
        t0 = time.perf_counter()
        u = isotropic_field(n, self.k_peak, self.u_rms, self.seed)
        self.log(f"[les] initial field {n}^3 in {time.perf_counter() - t0:.1f} s")
        g = periodic_grid(P, n, self.nu_fine, self.storage, u, self.cs_fine)
        del u
        nx, ny, _ = cases.tgv_shape(n)
        done = 0
        t_step = 0.0
        for s in range(0, steps + 1, every):
            if s > done:
                P.synchronize()
                t1 = time.perf_counter()
                P.updateGrid([g], s - done)
                P.synchronize()
                t_step += time.perf_counter() - t1
                done = s
            mac = g.macroscopic_device()
            N = nx * ny * n
            views = [_Offset(mac, k * N * 4) for k in range(4)]
            co = P.les_filter(*views, nx, ny, n, r)
            coarse = co.download((10, m, m, m))
            self.snapshots[s] = (coarse[0], coarse[1], coarse[2], coarse[3])
            if s >= first_sample:
                feat, tgt = P.les_targets(co, m, self.nu_coarse)
                self.feats.append(feat)
                self.tgts.append(tgt)
            del mac, co
            self.log(f"[les] fine step {s}: filtered snapshot ({'sampled' if s >= first_sample else 'reference'})")
        self.timings["fine_mlups"] = g.n * done / max(t_step, 1e-9) / 1e6
        g.close()

    def training_set(self):
        P, m = self.P, self.m
        k = len(self.feats)
        ns = m ** 3
        feat = P.DeviceBuffer(k * ns * 9 * 4)
        tgt = P.DeviceBuffer(k * ns * 4)
        for i, (f, t) in enumerate(zip(self.feats, self.tgts)):
            P.device_copy(_Offset(feat, i * ns * 9 * 4), f, ns * 9 * 4)
            P.device_copy(_Offset(tgt, i * ns * 4), t, ns * 4)
        return feat, tgt, k * ns

    def train(self, iters: int = 2000, batch: int = 16384, lr: float = 2e-3):
        P = self.P
        feat, tgt, ns = self.training_set()
        net = P.MLP(glorot_params(self.seed))
        t0 = time.perf_counter()
        hist = net.train(feat, tgt, ns, batch=batch, iters=iters, lr=lr, seed=self.seed)
        self.timings["train_s"] = time.perf_counter() - t0
        t = tgt.download(ns)
        self.target_stats = {"samples": ns, "mean": float(t.mean()), "std": float(t.std()),
                             "frac_zero": float(np.mean(t <= 0))}
        self.log(f"[les] trained on {ns} samples: loss {hist[0]:.4e} -> {hist[-1]:.4e} "
                 f"(target mean {t.mean():.3f}, std {t.std():.3f})")
        self.net, self.loss_hist = net, hist
        return net

    def coarse_run(self, model: str, start: int, param: float = 0.0, every: int = 1):
        """Coarse LES from the filtered snapshot at fine step `start`, spectra at
        every `every`-th later snapshot.  model "none", "smagorinsky" (param =
        Cs as in include/cellFunctions.h:192-197) or "mlp" (param = scale of the
        network output)."""
        P, m, r = self.P, self.m, self.r
        rho, ux, uy, uz = self.snapshots[start]
        g = periodic_grid(P, m, self.nu_coarse, "f32", (ux, uy, uz), param if model == "smagorinsky" else 0.0, rho)
        net = None
        if model == "mlp":
            net = self.scaled_net(param)
            g.set_closure(net)
            g.set_closure_every(self.closure_every)
        later = sorted(k for k in self.snapshots if k > start)[every - 1::every]
        errs = {}
        done = start
        for s in later:
            P.updateGrid([g], (s - done) // r)
            done = s
            e = energy_spectrum(*unique_velocity(g, m))
            e_box, e_dec = self.reference_spectrum(s)
            errs[s] = (spectrum_error(e, e_dec, m // 4), spectrum_error(e, e_dec, m // 2),
                       spectrum_error(e, e_box, m // 4))
        g.close()
        return errs

    def scaled_net(self, alpha: float):
        """The trained network with its output scaled by alpha (last layer)."""
        p, mean, std = self.net.get_params()
        q = p.copy()
        q[-33:] *= np.float32(alpha)  # W3 and b3
        return self.P.MLP(q, mean, std)

    def reference_spectrum(self, s: int):
        """Spectra of the filtered fine field at fine step s (cached): as
        filtered, and with the box filter's transfer divided out (the fine
        run's own spectrum at the coarse grid's scales, which a coarse LES
        should reproduce)."""
        if not hasattr(self, "_eref"):
            self._eref = {}
        if s not in self._eref:
            _, fx, fy, fz = self.snapshots[s]
            u = [a.astype(np.float64) for a in (fx, fy, fz)]
            self._eref[s] = (energy_spectrum(*u), energy_spectrum(*u, weight=1.0 / box_transfer2(self.m, self.r)))
        return self._eref[s]

    def compare(self, start: int, cs_values=(0.1, 0.17, 0.05, 0.02), alpha_values=(1.0, 2.0, 0.5, 0.25),
                every: int = 1, cs_reference: float = 0.1):
        """Coarse LES runs judged by their energy spectra against the fine run
        (box transfer divided out of the filtered field's spectrum) over the
        second half of the later snapshots, once the coarse runs have left the
        filtered start behind.  The primary figures are the models as they come:
        Smagorinsky at cs_reference (the reference drivers' constant,
        SmagorinskyConstantGlobal = 0.1) and the network as trained (output
        scale 1).  Each model is also run over a small parameter grid and its
        best value on the same window is reported beside, as an upper bound (a
        setting tuned on the data it is judged on).  Errors for shells k <= m/4
        (the resolved band, the primary figure), k <= m/2, and against the
        filtered field's own spectrum (k <= m/4); the first half of the window
        is reported as "early"."""
        runs = {("none", 0.0): self.coarse_run("none", start, 0.0, every)}
        for c in cs_values:
            runs[("smagorinsky", c)] = self.coarse_run("smagorinsky", start, c, every)
        for a in alpha_values:
            runs[("mlp", a)] = self.coarse_run("mlp", start, a, every)
        steps = sorted(next(iter(runs.values())).keys())
        half = len(steps) // 2
        early, ev = steps[:half], steps[half:]

        def score(errs, which, band):
            return float(np.mean([errs[s][band] for s in which]))

        def scores(errs):
            return {"evaluation_resolved": score(errs, ev, 0), "evaluation_all": score(errs, ev, 1),
                    "evaluation_vs_filtered": score(errs, ev, 2), "early": score(errs, early, 0)}

        out = {"early_steps": early, "evaluation_steps": ev, "runs": {}}
        for (model, prm), errs in runs.items():
            out["runs"][f"{model}:{prm}"] = scores(errs)
        out["none"] = scores(runs[("none", 0.0)])
        out["smagorinsky"] = dict(scores(runs[("smagorinsky", cs_reference)]), param=cs_reference)
        out["mlp"] = dict(scores(runs[("mlp", 1.0)]), param=1.0)
        for model in ("smagorinsky", "mlp"):
            cands = [(prm, e) for (mm, prm), e in runs.items() if mm == model]
            prm, errs = min(cands, key=lambda c: score(c[1], ev, 0))
            out[f"{model}_tuned"] = dict(scores(errs), param=prm)
        r = {k: out[k]["evaluation_resolved"] for k in ("none", "smagorinsky", "mlp", "smagorinsky_tuned", "mlp_tuned")}
        out["gain_vs_smagorinsky"] = 1.0 - r["mlp"] / r["smagorinsky"]
        out["gain_vs_none"] = 1.0 - r["mlp"] / r["none"]
        out["gain_vs_smagorinsky_tuned"] = 1.0 - r["mlp"] / r["smagorinsky_tuned"]
        out["gain_tuned_vs_tuned"] = 1.0 - r["mlp_tuned"] / r["smagorinsky_tuned"]
        out["closure_every"] = self.closure_every
        for k in ("none", "smagorinsky", "mlp", "smagorinsky_tuned", "mlp_tuned"):
            self.log(f"[les] {k}: resolved-band spectrum error {out[k]['evaluation_resolved']:.4f}"
                     + (f" (param {out[k]['param']})" if "param" in out[k] else ""))
        return out

    def cost_ratio(self, n: int = 256, launches: int = 96, storage: str = "f32", cs: float = 0.17,
                   repeats: int = 3):
        """Mean bulk-kernel time with the network / with Smagorinsky on a periodic
        n^3 box (launches back to back between two events; with the network
        evaluated every closure_every steps, the mean over evaluating and
        holding launches).  The two are timed alternately `repeats` times and
        each keeps its fastest, so that a slow spell of the device does not
        fall on one of them only."""
        P = self.P
        u = [np.zeros((n, n, n), np.float32)] * 3
        grids = {}
        for model in ("smagorinsky", "mlp"):
            g = periodic_grid(P, n, self.nu_coarse, storage, u, cs)
            if model == "mlp":
                g.set_closure(self.net)
                g.set_closure_every(self.closure_every)
            P.updateGrid([g], 2)
            grids[model] = g
        times = {m: float("inf") for m in grids}
        for _ in range(repeats):
            for model, g in grids.items():
                times[model] = min(times[model], g.time_update_kernel(launches)[0])
        for g in grids.values():
            g.close()
        return times["mlp"] / times["smagorinsky"], times


class _Offset:
    """A view of a DeviceBuffer at a byte offset (for the C ABI)."""

    def __init__(self, buf, offset: int):
        import ctypes

        self._buf = buf
        self.nbytes = buf.nbytes - offset
        self.ptr = ctypes.c_void_p(buf.ptr.value + offset)


# ------------------------------------------------- sphere at Re = 1e5 with the closure
def sphere_grid(P, scale: int, storage: str, reynolds: float, cs: float, u_inlet: float = 0.04):
    """The sphere channel (256 x 128 x 128 cells, D = 32, flowAroundSphereMainSingle.cu's
    layout: velocity inlet, pressure outlet, periodic sides) refined `scale` times:
    cells of size 1/scale nested in the unrefined cells, the same icosphere
    voxelized, the same lattice inlet velocity and so a lattice viscosity `scale`
    times larger at the same Reynolds number.  dtPhys = res / 1000 makes the
    section cuts report lattice velocities.  Starts from uniform flow."""
    nx, ny, nz = (a * scale for a in cases.SPHERE_SHAPE)
    res = 1.0 / scale
    off = -0.5 + 0.5 * res
    d = cases.SPHERE_DIAMETER * scale
    info = P.InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=u_inlet * d / reynolds, res=res,
                        ox=off, oy=off, oz=off, dtPhys=res / 1000.0)
    stl = P.STLStruct.from_triangles(cases.icosphere(4, cases.SPHERE_DIAMETER / 2, cases.SPHERE_CENTER))
    mask = P.BoolArray(nx * ny * nz)
    P.applyMarkersInsideSTL(mask, stl, True, info)
    solid = mask.get().astype(bool)
    stl.close()
    g = P.GridStruct(info, storage)
    g.set_flags(cases.channel_flags((nx, ny, nz), solid))
    g.set_given_values(cases.channel_given_values(u_inlet))
    g.set_smagorinsky(cs)
    g.set_periodic_copies(y=True)
    P.fill_equilibrium(g, ux=np.where(solid, 0.0, u_inlet).astype(np.float32))
    return g, d


def _sphere_run(P, g, scale: int, d: float, steps: int, every: int, plane_every: int, log, label: str):
    """Drag coefficient every `every` steps and the mid-plane (XY through the
    sphere centre) streamwise velocity every `plane_every` steps over the
    second half, averaged over the scale layers that make up the unrefined
    mid-plane cell and over scale x scale in-plane blocks.  Stops at the first
    sample with a non-finite or non-positive-density cell."""
    lo, hi = cases.sphere_box()
    lo = [a * scale for a in lo]
    hi = [a * scale for a in hi]
    k0 = int(cases.SPHERE_CENTER[2]) * scale
    cds, planes, solid = [], [], None
    t0 = time.perf_counter()
    for s in range(every, steps + 1, every):
        P.updateGrid([g], every)
        if s % (10 * every) == 0 or s == steps:
            bad = g.check_health()[0]
            if bad:
                log(f"[les-sphere] {label}: {bad} unhealthy cells after {s} steps")
                return {"diverged_at": s, "cd_history": cds}
        cds.append(cases.drag_coefficient(float(g.momentum_exchange(lo, hi)[0]), 0.04, d))
        if s > steps // 2 and s % plane_every == 0:
            cut = sum(P.section_cut(g, "XY", k0 + k) for k in range(scale)) / scale
            ny, nx = cut.shape[0] // scale, cut.shape[1] // scale
            blocks = cut.reshape(ny, scale, nx, scale, 5)
            planes.append(blocks[..., 1].mean(axis=(1, 3)))
            if solid is None:
                solid = blocks[..., 4].max(axis=(1, 3)) > 0
    el = time.perf_counter() - t0
    half = cds[len(cds) // 2:]
    log(f"[les-sphere] {label}: mean C_D {np.mean(half):.4f} over steps {steps // 2}-{steps} ({el:.1f} s)")
    return {"cd_mean": float(np.mean(half)), "cd_std": float(np.std(half)), "cd_history": cds,
            "ux_mean_plane": np.mean(planes, axis=0), "solid": solid, "seconds": el}


def sphere_comparison(P, net, coarse_steps: int = 20000, closure_every: int = 8, cs: float = 0.1,
                      reynolds: float = 1e5, storage: str = "f16", log=print):
    """The Re = 1e5 sphere on the 256 x 128 x 128 grid with no SGS model, with
    Smagorinsky (Cs = 0.1, the reference drivers' value) and with the learned
    closure `net` (evaluated every `closure_every` steps, box-face cells on
    Smagorinsky), judged against the same sphere on a 2x finer grid with
    Smagorinsky over the same physical time (twice the steps).  Errors: mean
    drag coefficient over the second half relative to the fine run's, and the
    relative L2 error of the time-averaged mid-plane streamwise velocity in the
    wake window (sphere centre to 4 D behind it, 1.5 D either side of the axis;
    fluid cells of both grids) against the fine run's, averaged onto the
    coarse cells."""
    every = 100
    plane_every = max(every, coarse_steps // 40 // every * every)  # 20 planes over the second half
    runs = {}
    g, d = sphere_grid(P, 2, storage, reynolds, cs)
    runs["fine"] = _sphere_run(P, g, 2, d, 2 * coarse_steps, 2 * every, 2 * plane_every, log,
                               "fine (2x, Smagorinsky)")
    g.close()
    for name in ("none", "smagorinsky", "mlp"):
        g, d = sphere_grid(P, 1, storage, reynolds, 0.0 if name == "none" else cs)
        if name == "mlp":
            g.set_closure(net)
            g.set_closure_every(closure_every)
        runs[name] = _sphere_run(P, g, 1, d, coarse_steps, every, plane_every, log, name)
        g.close()
    fine = runs["fine"]
    cx, cy = int(cases.SPHERE_CENTER[0]), int(cases.SPHERE_CENTER[1])
    D = int(cases.SPHERE_DIAMETER)
    win = (slice(cy - 3 * D // 2, cy + 3 * D // 2), slice(cx, cx + 4 * D))
    out = {"reynolds": reynolds, "coarse_steps": coarse_steps, "fine_steps": 2 * coarse_steps, "cs": cs,
           "closure_every": closure_every, "storage": storage,
           "fine": {k: fine[k] for k in ("cd_mean", "cd_std", "seconds") if k in fine}}
    for name in ("none", "smagorinsky", "mlp"):
        r = runs[name]
        if "diverged_at" in r or "diverged_at" in fine:
            out[name] = {"diverged_at": r.get("diverged_at"), "fine_diverged_at": fine.get("diverged_at")}
            continue
        keep = ~(r["solid"] | fine["solid"])[win]
        a, b = r["ux_mean_plane"][win][keep], fine["ux_mean_plane"][win][keep]
        out[name] = {"cd_mean": r["cd_mean"], "cd_std": r["cd_std"],
                     "cd_rel_err": abs(r["cd_mean"] / fine["cd_mean"] - 1),
                     "wake_ux_rel_l2": float(np.linalg.norm(a - b) / np.linalg.norm(b)), "seconds": r["seconds"]}
    return out
