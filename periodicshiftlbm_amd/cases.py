"""Workload builders: the driver-side callbacks of the benchmark configurations.

These play the role of the reference drivers' getMarkers / getGivenRhoUxUyUz /
getInitialRhoUxUyUz callbacks (flowAroundSphere/flowAroundSphereMainSingle.cu:41-92),
evaluated on the host into flag bytes and initial fields.

* Taylor-Green vortex on a fully periodic 64^3 box: the reference realises x/y
  periodicity with a duplicated boundary layer and applyBCPeriodicX/Y copies
  (include/boundaryConditions/periodicXY.h:28-31), so the grid is 65 x 65 x 64
  with period 64 in every direction (z is periodic through the shift itself).
* Flow past a sphere in a 256 x 128 x 128 channel: icosphere of diameter 32
  voxelized through the STL path, velocity inlet (given u) at i = 0, pressure
  outlet (given rho) at i = Nx-1, periodic y (copies) and z (implicit).
* A batch of independent sphere-in-channel domains of uneven size (2L x L x L,
  L uniform in [32, 96], sphere diameter L/4, Re log-uniform in [100, 5000])
  drawn from a counter-based generator with a fixed seed.  No reference
  counterpart: flowAroundSphere/flowAroundSphereMainBatch.cu:179-292 sweeps
  Reynolds numbers one grid at a time.
"""
from __future__ import annotations

import math

import numpy as np

FLUID, BOUNCEBACK, GIVEN_RHO, GIVEN_U, GIVEN_NONE, GIVEN_RHOU, GHOST, COVERED = range(8)
PERIODIC_X, PERIODIC_Y, PERIODIC_Z = 8, 16, 32


# ------------------------------------------------------------ Taylor-Green
def tgv_shape(n: int = 64):
    return (n + 1, n + 1, n)


def tgv_fields(n: int = 64, u0: float = 0.05):
    """rho, ux, uy, uz (flattened, cell = k*Nx*Ny + j*Nx + i) on the (n+1, n+1, n) grid."""
    nx, ny, nz = tgv_shape(n)
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    x = 2 * np.pi * i / n
    y = 2 * np.pi * j / n
    z = 2 * np.pi * k / n
    ux = u0 * np.sin(x) * np.cos(y) * np.cos(z)
    uy = -u0 * np.cos(x) * np.sin(y) * np.cos(z)
    uz = np.zeros_like(ux)
    p = (u0 * u0 / 16.0) * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2.0)
    rho = 1.0 + 3.0 * p
    return [a.reshape(-1) for a in (rho, ux, uy, uz)]


def kinetic_energy_enstrophy(rho, ux, uy, uz, n: int = 64):
    """Kinetic energy 1/2 sum rho |u|^2 and enstrophy 1/2 sum |curl u|^2 (central
    differences) over the n^3 unique cells of the (n+1, n+1, n) periodic grid."""
    nx, ny, nz = tgv_shape(n)
    sh = (nz, ny, nx)
    r, a, b, c = (np.asarray(v, np.float64).reshape(sh)[:, :n, :n] for v in (rho, ux, uy, uz))
    ke = 0.5 * float(np.sum(r * (a * a + b * b + c * c)))

    def d(f, axis):
        return 0.5 * (np.roll(f, -1, axis) - np.roll(f, 1, axis))

    # axes: 0 = z, 1 = y, 2 = x
    wx = d(c, 1) - d(b, 0)
    wy = d(a, 0) - d(c, 2)
    wz = d(b, 2) - d(a, 1)
    ens = 0.5 * float(np.sum(wx * wx + wy * wy + wz * wz))
    return ke, ens


# ------------------------------------------------------------ icosphere
def icosphere(subdivisions: int = 4, radius: float = 16.0, center=(64.0, 64.0, 64.0)) -> np.ndarray:
    """Triangles (T, 9) of an icosphere with outward orientation (20 * 4^s triangles)."""
    t = (1.0 + 5 ** 0.5) / 2.0
    verts = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
             (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    verts = [np.array(v, np.float64) / np.linalg.norm(v) for v in verts]
    faces = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2),
             (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11),
             (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    for _ in range(subdivisions):
        cache = {}

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = verts[a] + verts[b]
                verts.append(m / np.linalg.norm(m))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius + np.array(center)
    # vertices as float32 first so shared edges compare equal exactly
    V32 = V.astype(np.float32)
    tri = np.stack([V32[[f[0] for f in faces]], V32[[f[1] for f in faces]], V32[[f[2] for f in faces]]], axis=1)
    return tri.reshape(-1, 9)


def write_binary_stl(path: str, tri: np.ndarray):
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1, 9)
    rec = np.zeros(tri.shape[0], dtype=[("n", "<f4", 3), ("v", "<f4", 9), ("a", "<u2")])
    rec["v"] = tri
    with open(path, "wb") as fh:
        fh.write(b"periodicshiftlbm_amd icosphere".ljust(80, b" "))
        fh.write(np.uint32(tri.shape[0]).tobytes())
        fh.write(rec.tobytes())


# ------------------------------------------------------------ sphere channel
SPHERE_SHAPE = (256, 128, 128)
SPHERE_CENTER = (64.0, 64.0, 64.0)
SPHERE_DIAMETER = 32.0


def sphere_nu(u_inlet: float = 0.04, diameter: float = SPHERE_DIAMETER, reynolds: float = 1000.0) -> float:
    return u_inlet * diameter / reynolds


def channel_flags(shape, solid: np.ndarray | None = None) -> np.ndarray:
    """Flag bytes of the sphere channel (cell = k*Nx*Ny + j*Nx + i).

    slot 0 = inlet (given u), slot 1 = outlet (given rho); y/z faces are periodic
    (the inlet/outlet cells on those faces keep the x normal only)."""
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    flags = np.full((nz, ny, nx), FLUID, np.uint8)
    per = PERIODIC_Y | PERIODIC_Z
    flags[i == 0] = GIVEN_U | per | (0 << 6)
    flags[i == nx - 1] = GIVEN_RHO | per | (1 << 6)
    flags = flags.reshape(-1)
    if solid is not None:
        s = np.asarray(solid, bool).reshape(-1)
        flags[s] = BOUNCEBACK
    return flags


def channel_given_values(u_inlet: float = 0.04):
    return np.array([[1.0, u_inlet, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0], [1.0, 0.0, 0.0, 0.0]],
                    np.float32)


def sphere_box(center=SPHERE_CENTER, diameter=SPHERE_DIAMETER):
    """Bounding box [lo, hi) of the solid for the drag reduction
    (flowAroundSphereMainSingle.cu:105-110)."""
    r = diameter / 2
    lo = [int(c - r) for c in center]
    hi = [int(c + r) + 2 for c in center]
    return lo, hi


def drag_coefficient(fx: float, u_inlet: float = 0.04, diameter: float = SPHERE_DIAMETER) -> float:
    """C_D = F / (1/2 rho U^2 pi D^2 / 4) with the momentum-exchange force in lattice units;
    the reference's sign convention gives C_D = -8 F / (rho U^2 pi D^2) for F = sum rho (u - u_bb)
    (flowAroundSphereMainSingle.cu:250)."""
    return -8.0 * fx / (u_inlet * u_inlet * math.pi * diameter * diameter)


# ------------------------------------------------------------ domain batch
_MASK64 = (1 << 64) - 1


def counter_uniform(seed: int, counter: int) -> float:
    """Uniform [0, 1) from (seed, counter): the splitmix64 finaliser applied to
    seed + (counter + 1) * 0x9E3779B97F4A7C15, top 53 bits."""
    z = (seed + (counter + 1) * 0x9E3779B97F4A7C15) & _MASK64
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _MASK64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _MASK64
    z ^= z >> 31
    return (z >> 11) / float(1 << 53)


def batch_domains(count: int = 64, seed: int = 20240607, l_min: int = 32, l_max: int = 96,
                  re_min: float = 100.0, re_max: float = 5000.0, u_inlet: float = 0.04):
    """Parameters of `count` sphere-in-channel domains: edge L uniform in
    [l_min, l_max], box 2L x L x L, sphere of diameter L/4 centred at
    (L/2, L/2, L/2), Re log-uniform in [re_min, re_max], nu = U D / Re."""
    out = []
    for d in range(count):
        L = l_min + int(counter_uniform(seed, 2 * d) * (l_max - l_min + 1))
        re = re_min * (re_max / re_min) ** counter_uniform(seed, 2 * d + 1)
        diameter = 0.25 * L
        out.append({"L": L, "shape": (2 * L, L, L), "re": re, "diameter": diameter,
                    "center": (0.5 * L, 0.5 * L, 0.5 * L), "nu": u_inlet * diameter / re, "u_inlet": u_inlet})
    return out


def walled_channel_flags(shape, solid: np.ndarray | None = None) -> np.ndarray:
    """Flag bytes of a batch domain: velocity inlet (slot 0) at i = 0, pressure
    outlet (slot 1) at i = Nx-1, bounceback walls at j = 0 and j = Ny-1, z
    periodic through the shift (the inlet/outlet normal ignores z)."""
    nx, ny, nz = shape
    k, j, i = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    flags = np.full((nz, ny, nx), FLUID, np.uint8)
    flags[i == 0] = GIVEN_U | PERIODIC_Z | (0 << 6)
    flags[i == nx - 1] = GIVEN_RHO | PERIODIC_Z | (1 << 6)
    flags[(j == 0) | (j == ny - 1)] = BOUNCEBACK
    flags = flags.reshape(-1)
    if solid is not None:
        flags[np.asarray(solid, bool).reshape(-1)] = BOUNCEBACK
    return flags


def sphere_box_of(center, diameter):
    """[lo, hi) box around a sphere for the momentum-exchange reduction."""
    r = diameter / 2
    return [int(c - r) - 1 for c in center], [int(c + r) + 2 for c in center]


def domain_grid(dom: dict, storage: str = "f16", cs: float = 0.17, subdivisions: int = 3):
    """GridStruct of one batch domain: the sphere is an icosphere voxelized by
    the STL path, the grid starts at rest (rho = 1, u = 0)."""
    from . import (BoolArray, GridStruct, InfoStruct, STLStruct, applyMarkersInsideSTL, fill_equilibrium)

    nx, ny, nz = dom["shape"]
    info = InfoStruct(cellCountX=nx, cellCountY=ny, cellCountZ=nz, nu=dom["nu"], res=1.0)
    stl = STLStruct.from_triangles(icosphere(subdivisions, dom["diameter"] / 2, dom["center"]))
    mask = BoolArray(nx * ny * nz)
    applyMarkersInsideSTL(mask, stl, True, info)
    flags = walled_channel_flags(dom["shape"], mask.get().astype(bool))
    g = GridStruct(info, storage)
    g.set_flags(flags)
    g.set_given_values(channel_given_values(dom["u_inlet"]))
    g.set_smagorinsky(cs)
    fill_equilibrium(g)
    return g, flags


# ------------------------------------------------------------ two-level sphere
# Fine box of the two-level sphere configuration in coarse cells [start, end).
TWO_LEVEL_SUB = (32, 16, 16, 160, 112, 112)


def two_level_sphere(storage: str = "f16", reynolds: float = 1e5, u_inlet: float = 0.04, cs: float = 0.17,
                     sub=TWO_LEVEL_SUB, shape=None, center=None, diameter=None):
    """The sphere channel (256 x 128 x 128 coarse cells) with a 2x finer nested
    box around the sphere (flowAroundSphere/flowAroundSphereMainSingle.cu:188-213
    sets sub-grid bounds and the fine origin ox_c + iStart res_c - res_f / 2).
    The fine grid has ghost box faces; the sphere is voxelized on both levels
    by the STL path.  shape / center / diameter (coarse cells) default to the
    256 x 128 x 128 channel.  Returns ([coarse, fine], fine sphere box (lo, hi))."""
    from . import BoolArray, GridStruct, InfoStruct, STLStruct, applyMarkersInsideSTL, fill_equilibrium

    center = center or SPHERE_CENTER
    diameter = diameter or SPHERE_DIAMETER
    nxc, nyc, nzc = shape or SPHERE_SHAPE
    nu_c = u_inlet * diameter / reynolds
    ci = InfoStruct(cellCountX=nxc, cellCountY=nyc, cellCountZ=nzc, nu=nu_c, res=1.0,
                    iSubgridStart=sub[0], jSubgridStart=sub[1], kSubgridStart=sub[2],
                    iSubgridEnd=sub[3], jSubgridEnd=sub[4], kSubgridEnd=sub[5])
    nxf, nyf, nzf = 2 * (sub[3] - sub[0]), 2 * (sub[4] - sub[1]), 2 * (sub[5] - sub[2])
    ox, oy, oz = sub[0] - 0.25, sub[1] - 0.25, sub[2] - 0.25
    fi = InfoStruct(gridID=1, cellCountX=nxf, cellCountY=nyf, cellCountZ=nzf, nu=2 * nu_c, res=0.5,
                    ox=ox, oy=oy, oz=oz)
    stl = STLStruct.from_triangles(icosphere(4, diameter / 2, center))
    grids = []
    for info, fine in ((ci, False), (fi, True)):
        n = info.cellCountX * info.cellCountY * info.cellCountZ
        mask = BoolArray(n)
        applyMarkersInsideSTL(mask, stl, True, info)
        solid = mask.get().astype(bool)
        dims = (info.cellCountX, info.cellCountY, info.cellCountZ)
        if fine:
            k, j, i = np.meshgrid(np.arange(dims[2]), np.arange(dims[1]), np.arange(dims[0]), indexing="ij")
            ghost = ((i == 0) | (i == dims[0] - 1) | (j == 0) | (j == dims[1] - 1) | (k == 0)
                     | (k == dims[2] - 1)).reshape(-1)
            flags = np.where(ghost, GHOST, np.where(solid, BOUNCEBACK, FLUID)).astype(np.uint8)
        else:
            flags = channel_flags(dims, solid)
        g = GridStruct(info, storage)
        g.set_flags(flags)
        g.set_given_values(channel_given_values(u_inlet))
        g.set_smagorinsky(cs)
        if not fine:
            g.set_periodic_copies(y=True)
        fill_equilibrium(g)
        grids.append(g)
    # sphere box in fine cell indices (cell i centre at ox + (i + 1/2) res_f)
    r = diameter / 2
    lo = [int((c - r - o) / 0.5) - 2 for c, o in zip(center, (ox, oy, oz))]
    hi = [int((c + r - o) / 0.5) + 3 for c, o in zip(center, (ox, oy, oz))]
    return grids, (lo, hi)


# ------------------------------------------------ nested periodic-shift sphere
def reference_sphere_levels(levels: int = 4, cells: int = 220, reynolds: float = 1e6, u_inlet: float = 0.07):
    """The nested-box sphere of flowAroundSphere/flowAroundSphereMainSingle.cu:
    a sphere of 2000 mm at (4000, 11000, 11000) mm in a cube of 22000 mm with
    `cells` coarse cells per edge (220: the driver's 100 mm), and `levels`
    grids, each 2x finer, nested by the driver's placement rule (lines
    186-213: progress = (l / (L-1))^0.2 sets the box, the fine origin is
    ox_c + iStart res_c - res_f / 2).  Markers (lines 41-67): bounceback inside
    the sphere (r^2 <= R^2 at iCell * res + ox), given u on the coarse
    grid's j, k and i = 0 faces, given rho at i = Nx - 1; ghost box faces on
    the finer grids.  Every expression is evaluated in FP32 in the driver's
    order.  Returns a list of (InfoStruct fields dict, flag bytes) and the
    finest level's drag box (lo, hi) of getSphereDrag (lines 99-112)."""
    f32 = np.float32
    D = f32(2000.0)
    R = f32(0.5 * 2000.0)
    res0 = f32(22000.0 / cells)
    ux = f32(u_inlet)
    nu_phys = f32(ux * (D / f32(1000.0)) / f32(reynolds))
    dt0 = f32(f32(ux / ux) * f32(res0 / f32(1000)))
    domain = f32(f32(11.0) * D)
    sx, sy, sz = f32(2.0) * D, f32(5.5) * D, f32(5.5) * D
    n0 = int(math.ceil(float(domain / res0)))

    def nu_of(dt, res):
        r = f32(res / f32(1000.0))
        return f32(f32(dt * nu_phys) / f32(r * r))

    infos = [dict(gridID=0, res=res0, dtPhys=dt0, nu=nu_of(dt0, res0), cellCountX=n0, cellCountY=n0, cellCountZ=n0,
                  ox=f32(0), oy=f32(0), oz=f32(0))]
    for lev in range(1, levels):
        c = infos[-1]
        res = f32(c["res"] * f32(0.5))
        dt = f32(c["dtPhys"] * f32(0.5))
        progress = f32(np.power(f32(f32(lev) / f32(levels - 1)), f32(0.2)))
        x_start = f32(f32(progress * f32(1.4)) * D)
        x_end = f32(f32(f32(1) - progress) * domain) + f32(f32(progress * f32(2.6)) * D)
        y_start = f32(f32(progress * f32(4.9)) * D)
        c["iSubgridStart"] = int(f32(f32(f32(x_start - c["ox"]) / c["res"]) + f32(0.5)))
        c["iSubgridEnd"] = int(f32(f32(f32(x_end - c["ox"]) / c["res"]) + f32(0.5)))
        c["jSubgridStart"] = int(f32(f32(f32(y_start - c["oy"]) / c["res"]) + f32(0.5)))
        c["jSubgridEnd"] = c["jSubgridStart"] + (c["iSubgridEnd"] - c["iSubgridStart"])
        c["kSubgridStart"] = c["jSubgridStart"]
        c["kSubgridEnd"] = c["jSubgridEnd"]
        half = f32(res * f32(0.5))
        infos.append(dict(
            gridID=lev, res=res, dtPhys=dt, nu=nu_of(dt, res),
            ox=f32(f32(c["ox"] + f32(f32(c["iSubgridStart"]) * c["res"])) - half),
            oy=f32(f32(c["oy"] + f32(f32(c["jSubgridStart"]) * c["res"])) - half),
            oz=f32(f32(c["oz"] + f32(f32(c["kSubgridStart"]) * c["res"])) - half),
            cellCountX=2 * (c["iSubgridEnd"] - c["iSubgridStart"]),
            cellCountY=2 * (c["jSubgridEnd"] - c["jSubgridStart"]),
            cellCountZ=2 * (c["kSubgridEnd"] - c["kSubgridStart"])))
    out = []
    for info in infos:
        nx, ny, nz = info["cellCountX"], info["cellCountY"], info["cellCountZ"]
        k, j, i = np.meshgrid(np.arange(nz, dtype=np.int32), np.arange(ny, dtype=np.int32),
                              np.arange(nx, dtype=np.int32), indexing="ij")
        x = i.astype(f32) * info["res"] + info["ox"]
        y = j.astype(f32) * info["res"] + info["oy"]
        z = k.astype(f32) * info["res"] + info["oz"]
        r2 = (x - sx) * (x - sx) + (y - sy) * (y - sy) + (z - sz) * (z - sz)
        solid = r2 <= R * R
        if info["gridID"] != 0:
            face = (i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)
            flags = np.where(solid, BOUNCEBACK, np.where(face, GHOST, FLUID))
        else:
            flags = np.full(solid.shape, FLUID)
            flags[i == nx - 1] = GIVEN_RHO
            flags[i == 0] = GIVEN_U
            flags[(k == 0) | (k == nz - 1)] = GIVEN_U
            flags[(j == 0) | (j == ny - 1)] = GIVEN_U
            flags[solid] = BOUNCEBACK
        out.append((info, flags.astype(np.uint8).reshape(-1)))
    fin = infos[-1]
    lo = [int(f32(f32(sx - R - fin["ox"]) / fin["res"])), int(f32(f32(sy - R - fin["oy"]) / fin["res"])),
          int(f32(f32(sz - R - fin["oz"]) / fin["res"]))]
    hi = [int(f32(f32(sx + R - fin["ox"]) / fin["res"])) + 2, int(f32(f32(sy + R - fin["oy"]) / fin["res"])) + 2,
          int(f32(f32(sz + R - fin["oz"]) / fin["res"])) + 2]
    return out, (lo, hi)


def reference_sphere_drag_coefficient(fx: float, levels_info, u_inlet: float = 0.07) -> float:
    """C_D of the finest level's lattice force: -8 F / (U^2 pi D^2) with D in
    finest cells (the driver's physical form reduces to this because its
    physical velocity equals the lattice velocity)."""
    d = 2000.0 / float(levels_info[-1]["res"])
    return -8.0 * fx / (u_inlet * u_inlet * math.pi * d * d)


# ------------------------------------------------------------ DIAD sphere
def diad_sphere_flags(ijk: np.ndarray, bounceback: np.ndarray, shape, level: int) -> np.ndarray:
    """Flag bytes of one level of the wall-refined sphere channel
    (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:41-96): on level 0
    given u on the j, k and i = 0 faces, given rho at i = Nx-1; on finer levels
    the box faces are ghost cells.  Bounceback cells come from the builder."""
    i, j, k = (np.asarray(a) for a in ijk)
    nx, ny, nz = shape
    flags = np.full(i.size, FLUID, np.uint8)
    if level == 0:
        flags[i == nx - 1] = GIVEN_RHO | (1 << 6)
        flags[i == 0] = GIVEN_U
        flags[(k == 0) | (k == nz - 1)] = GIVEN_U
        flags[(j == 0) | (j == ny - 1)] = GIVEN_U
    else:
        edge = (i == 0) | (i == nx - 1) | (j == 0) | (j == ny - 1) | (k == 0) | (k == nz - 1)
        flags[edge] = GHOST
    flags[np.asarray(bounceback, bool)] = BOUNCEBACK
    return flags


def diad_sphere_rules(P, radius: float, levels: int, refine=((1.5, 4.0), (1.0, 2.5), (0.7, 1.0), (0.6, 0.6))):
    """Marker rules of the DIAD sphere driver: the sphere wall on every level and,
    on the coarser levels, a refinement region made of a wake cylinder of
    half-width w D and length l D downstream plus a sphere of radius w D
    (flowAroundSphereMainSingleDIAD.cu:47-80)."""
    d = 2.0 * radius
    rs = [P.marker_rule(P.RULE_BOUNCEBACK, P.SHAPE_SPHERE_LE, p=(0.0, 0.0, 0.0, radius))]
    for lev, (w, l) in enumerate(refine[:max(0, levels - 1)]):
        rs.append(P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_XCYLINDER, p=(0.0, 0.0, 0.0, w * d, l * d), levels=[lev]))
        rs.append(P.marker_rule(P.RULE_REFINEMENT, P.SHAPE_SPHERE_LT, p=(0.0, 0.0, 0.0, w * d), levels=[lev]))
    return P.rules(*rs)


def diad_sphere(P, storage="f32", n0: int = 24, levels: int = 3, span: int = 1, radius: float = 2.0,
                res: float = 1.0, nu0: float = 0.02, u_inlet: float = 0.04, cs: float = 0.1):
    """Wall-refined sphere channel on DIAD levels: a base box of 2 n0 x n0 x n0
    cells of size res with the sphere (centre at the origin) at a quarter of the
    length; returns (levels, grids).  Physical viscosity and time step are chosen
    so that level 0 has lattice viscosity nu0 (finer levels: 2^l nu0)."""
    dt = 1e-3
    nu_phys = nu0 * (res / 1000.0) ** 2 / dt
    info = P.InfoStruct(res=res, cellCountX=2 * n0, cellCountY=n0, cellCountZ=n0, ox=-0.5 * n0 * res + 0.5 * res,
                        oy=-0.5 * n0 * res + 0.5 * res, oz=-0.5 * n0 * res + 0.5 * res, nu=nu0, dtPhys=dt)
    lv = P.buildDIADGrids(info, levels, span, nu_phys, (), diad_sphere_rules(P, radius, levels))
    grids = []
    for l in range(levels):
        g = P.DIADGridStruct.from_levels(lv, l, storage)
        li, _ = lv.info(l)
        g.set_flags(diad_sphere_flags(lv.array(l, "ijk"), lv.array(l, "bounceback"),
                                      (li.cellCountX, li.cellCountY, li.cellCountZ), l))
        g.set_given_values(channel_given_values(u_inlet))
        g.set_smagorinsky(cs)
        g.fill_equilibrium()
        grids.append(g)
    return lv, grids


def diad_sphere_reference(P, storage="f16", levels: int = 6, span: int = 3, u_inlet: float = 0.07,
                          reynolds: float = 1e6, cs: float = 0.1):
    """The wall-refined sphere of the DIAD sphere driver
    (flowAroundSphereDIAD/flowAroundSphereMainSingleDIAD.cu:1-29, 222-247):
    D = 2000 mm, base resolution 100 mm, an 11 D cube (220^3 base cells) with the
    sphere 2 D behind the inlet, `levels` levels with wall refinement span 3,
    u = 0.07, Re = 1e6, Smagorinsky 0.1, started at rest.  Returns (levels, grids)."""
    d, res = 2000.0, 100.0
    nu_phys = u_inlet * (d / 1000.0) / reynolds
    dt = (res / 1000.0)  # u_lattice = u_phys
    n = int(math.ceil(11.0 * d / res))
    info = P.InfoStruct(res=res, cellCountX=n, cellCountY=n, cellCountZ=n, ox=-2.0 * d, oy=-5.5 * d, oz=-5.5 * d,
                        nu=dt * nu_phys / (res / 1000.0) ** 2, dtPhys=dt)
    lv = P.buildDIADGrids(info, levels, span, nu_phys, (), diad_sphere_rules(P, d / 2, levels))
    grids = []
    for l in range(levels):
        g = P.DIADGridStruct.from_levels(lv, l, storage)
        li, _ = lv.info(l)
        g.set_flags(diad_sphere_flags(lv.array(l, "ijk"), lv.array(l, "bounceback"),
                                      (li.cellCountX, li.cellCountY, li.cellCountZ), l))
        g.set_given_values(channel_given_values(u_inlet))
        g.set_smagorinsky(cs)
        g.fill_equilibrium()
        grids.append(g)
    return lv, grids
