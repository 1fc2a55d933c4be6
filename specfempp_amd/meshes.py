"""Structured meshes and the benchmark workloads.

2-D meshes use the layout of the SPECFEM 2-D mesher's database (4-node
elements numbered x fastest; control nodes ordered bottom-left, bottom-right,
top-right, top-left as the shape functions of src/jacobian/shape_functions.cpp
expect; absorbing edges as (element, edge) pairs).  3-D meshes are the
structured hexahedral description of ``sfx_mesh3d``.

Time steps: the stable step of the NGLL = 5 scheme scales with the smallest
GLL spacing, 0.1727 h.  Workloads whose stated dt exceeds a Courant number of
0.5 on that spacing use dt = 0.4 * 0.1727 h / vp_max instead, and record both
(``dt_stated`` / ``dt``).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

from .io.fortran_db import EDGE_BOTTOM, EDGE_LEFT, EDGE_RIGHT, EDGE_TOP, Material, Mesh2D

GLL_MIN_SPACING = 0.5 * (1.0 - 0.6546536707079771)  # smallest GLL gap on [-1, 1] is 0.3453 -> 0.1727 h


def courant(dt: float, h: float, vp_max: float) -> float:
    return dt * vp_max / (GLL_MIN_SPACING * h)


def stable_dt(h: float, vp_max: float, c: float = 0.4) -> float:
    return c * GLL_MIN_SPACING * h / vp_max


def quad_mesh(nx: int, nz: int, xmin: float, xmax: float, zmin: float, zmax: float, material=None,
              materials=None, absorbing=("bottom", "right", "left")) -> Mesh2D:
    """Structured 4-node quad mesh; ``material`` is an (nz, nx) index array."""
    xs = np.linspace(xmin, xmax, nx + 1)
    zs = np.linspace(zmin, zmax, nz + 1)
    X, Z = np.meshgrid(xs, zs)  # (nz+1, nx+1)
    coorg = np.stack([X.ravel(), Z.ravel()])  # (2, npgeo), node = iz*(nx+1) + ix
    iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    n0 = (iz * (nx + 1) + ix).ravel()
    knods = np.stack([n0, n0 + 1, n0 + 1 + (nx + 1), n0 + (nx + 1)]).astype(np.int32)  # (4, nspec)
    if material is None:
        material = np.zeros((nz, nx), dtype=np.int32)
    kmato = np.ascontiguousarray(material, dtype=np.int32).ravel()
    # absorbing edges element-major, within an element in the order bottom,
    # right, top, left
    flags = set(absorbing)
    ez, ex = np.divmod(np.arange(nx * nz, dtype=np.int64), nx)
    on = np.stack([(ez == 0) & ("bottom" in flags), (ex == nx - 1) & ("right" in flags),
                   (ez == nz - 1) & ("top" in flags), (ex == 0) & ("left" in flags)], axis=1)
    codes = np.array([EDGE_BOTTOM, EDGE_RIGHT, EDGE_TOP, EDGE_LEFT], dtype=np.int32)
    elem, which = np.nonzero(on)
    if not isinstance(materials, MaterialArrays):
        materials = list(materials)
    return Mesh2D(nspec=nx * nz, npgeo=(nx + 1) * (nz + 1), ngnod=4, coorg=coorg, knods=knods, kmato=kmato,
                  materials=materials, abs_ispec=elem.astype(np.int32), abs_type=codes[which])


class MaterialArrays:
    """Isotropic elastic materials held as arrays (density, cp, cs), for
    meshes with a material per element; indexing yields a ``Material``."""

    def __init__(self, density, cp, cs):
        self.density = np.ascontiguousarray(density, dtype=np.float64).ravel()
        self.cp = np.ascontiguousarray(cp, dtype=np.float64).ravel()
        self.cs = np.ascontiguousarray(cs, dtype=np.float64).ravel()

    def __len__(self):
        return len(self.density)

    def __getitem__(self, i):
        return elastic(float(self.density[i]), float(self.cp[i]), float(self.cs[i]))


def material_arrays(materials):
    """(density, cp, cs) float64 arrays of a material list or ``MaterialArrays``."""
    if isinstance(materials, MaterialArrays):
        return materials.density, materials.cp, materials.cs
    return (np.array([m.density for m in materials], dtype=np.float64),
            np.array([m.cp for m in materials], dtype=np.float64),
            np.array([m.cs for m in materials], dtype=np.float64))


def elastic(rho: float, vp: float, vs: float) -> Material:
    return Material(indic=1, density=rho, cp=vp, cs=vs, qkappa=9999.0, qmu=9999.0,
                    raw=(rho, vp, vs, 0, 0, 9999.0, 9999.0, 0, 0, 0, 0, 0, 0))


@dataclass
class Workload2D:
    name: str
    mesh: Mesh2D
    sources: list
    receivers: list
    dt: float
    nstep: int
    dt_stated: float = 0.0
    notes: dict = field(default_factory=dict)


def config1(nstep: int = 1600, n: int = 80, h: float = 50.0) -> Workload2D:
    """2-D homogeneous P-SV: 80x80 quads of 50 m, free top, Stacey elsewhere,
    vertical Ricker force (10 Hz) at (2000, 2000), 50 receivers 500 m deep."""
    L = n * h
    mesh = quad_mesh(n, n, 0.0, L, 0.0, L, materials=[elastic(2200.0, 2500.0, 1443.0)],
                     absorbing=("bottom", "right", "left"))
    src = [dict(x=L / 2, z=L / 2, angle=0.0, f0=10.0, factor=1e10, tshift=0.0)]
    rec = [dict(x=40.0 + 80.0 * i * (L / 4000.0), z=L - 500.0 * (L / 4000.0)) for i in range(50)]
    return Workload2D("2d_homogeneous_80x80", mesh, src, rec, 1e-3, nstep, 1e-3,
                      {"courant": courant(1e-3, h, 2500.0)})


def gaussian_random_field(L: float, corr: float, seed, grid: float = None):
    """Zero-mean, unit-variance Gaussian random field with Gaussian covariance
    of correlation length corr on [0, L]^2: separable Gaussian filtering of
    N(0, 1) white noise on a lattice of spacing grid (default corr / 8),
    normalised, clipped at +-3 sigma.  Returns (axis, values[iz, ix])."""
    rng = np.random.default_rng(seed)
    grid = corr / 8.0 if grid is None else grid
    n = int(round(L / grid)) + 1
    x = np.arange(n) * grid
    white = rng.standard_normal((n, n))
    # a Gaussian covariance exp(-r^2 / corr^2) is white noise filtered by a
    # Gaussian kernel of standard deviation corr / 2
    sig = corr / 2.0 / grid
    half = int(np.ceil(4 * sig))
    t = np.arange(-half, half + 1)
    ker = np.exp(-0.5 * (t / sig) ** 2)
    ker /= ker.sum()
    f = np.apply_along_axis(lambda r: np.convolve(r, ker, mode="same"), 1, white)
    f = np.apply_along_axis(lambda c: np.convolve(c, ker, mode="same"), 0, f)
    f = (f - f.mean()) / f.std()
    return x, np.clip(f, -3.0, 3.0)


def config2(n: int = 4000, L: float = 10000.0, nstep: int = 4000, nrec: int = 200, seed: int = 2,
            corr: float = 200.0) -> Workload2D:
    """2-D heterogeneous P-SV: n x n quads over L x L (the full configuration
    is 4000^2 elements of 2.5 m), vp = 3000 (1 + 0.1 g) with g the Gaussian
    random field above (sampled at element centres: one material per element),
    vs = vp / sqrt(3), rho = 310 vp^0.25, Stacey on all sides, a vertical
    Ricker force (15 Hz) at a seeded point of the inner 80 %, nrec receivers on
    a seeded line; dt at Courant 0.4 (the stated 3e-4 s is unstable)."""
    h = L / n
    axis, g = gaussian_random_field(L, corr, seed)
    c = (np.arange(n) + 0.5) * h
    ix = np.clip(np.searchsorted(axis, c) - 1, 0, len(axis) - 2)
    tx = (c - axis[ix]) / (axis[ix + 1] - axis[ix])
    gx = g[:, ix] * (1 - tx) + g[:, ix + 1] * tx            # interpolate along x at every grid row
    gz = gx[ix, :] * (1 - tx)[:, None] + gx[ix + 1, :] * tx[:, None]   # then along z
    vp = 3000.0 * (1.0 + 0.1 * gz)                            # (nz, nx) per element
    vs = vp / np.sqrt(3.0)
    rho = 310.0 * vp ** 0.25
    mats = MaterialArrays(rho, vp, vs)
    material = np.arange(n * n, dtype=np.int32).reshape(n, n)
    mesh = quad_mesh(n, n, 0.0, L, 0.0, L, material, mats, ("bottom", "right", "left", "top"))
    rng = np.random.default_rng(seed)
    sx, sz = rng.uniform(0.1 * L, 0.9 * L, size=2)
    z_line = rng.uniform(0.2 * L, 0.8 * L)
    rec = [dict(x=float(x), z=float(z_line)) for x in np.sort(rng.uniform(0.05 * L, 0.95 * L, size=nrec))]
    src = [dict(x=float(sx), z=float(sz), angle=0.0, f0=15.0, factor=1e10, tshift=0.0)]
    dt = stable_dt(h, float(vp.max()))
    return Workload2D(f"2d_grf_{n}x{n}", mesh, src, rec, dt, nstep, 3e-4,
                      {"h": h, "vp_range": (float(vp.min()), float(vp.max())), "courant_stated": courant(3e-4, h, float(vp.max()))})


def config4(nsim: int = 64, seed: int = 4, nmin: int = 40, nmax: int = 240, h: float = 50.0,
            nstep_range=(800, 2400), nrec: int = 32):
    """Batched 2-D multi-event workload: nsim homogeneous simulations, side
    n ~ U{nmin..nmax} quads of h, vp ~ U[2000, 4000], vs = vp / sqrt(3),
    rho 2200, free top and Stacey elsewhere, one vertical Ricker force
    (f0 ~ U[5, 20] Hz) uniform in the inner 80 %, nstep ~ U{nstep_range},
    dt at Courant 0.4, nrec receivers uniform in the inner 80 %.  One
    PCG64 stream per simulation, spawned from the seed."""
    out = []
    for i, ss in enumerate(np.random.SeedSequence(seed).spawn(nsim)):
        rng = np.random.default_rng(ss)
        n = int(rng.integers(nmin, nmax + 1))
        vp = float(rng.uniform(2000.0, 4000.0))
        vs = vp / np.sqrt(3.0)
        f0 = float(rng.uniform(5.0, 20.0))
        nstep = int(rng.integers(nstep_range[0], nstep_range[1] + 1))
        L = n * h
        lo, hi = 0.1 * L, 0.9 * L
        sx, sz = rng.uniform(lo, hi, size=2)
        rx, rz = rng.uniform(lo, hi, size=(2, nrec))
        mesh = quad_mesh(n, n, 0.0, L, 0.0, L, materials=[elastic(2200.0, vp, vs)],
                         absorbing=("bottom", "right", "left"))
        src = [dict(x=float(sx), z=float(sz), angle=0.0, f0=f0, factor=1e10, tshift=0.0)]
        rec = [dict(x=float(x), z=float(z)) for x, z in zip(rx, rz)]
        out.append(Workload2D(f"batch{i}_{n}x{n}", mesh, src, rec, stable_dt(h, vp), nstep, stable_dt(h, vp),
                              {"n": n, "vp": vp, "f0": f0}))
    return out


def config5(n: int = 200, h: float = 20.0, nstep: int = 3000, dt: float = 5e-4, nsrc: int = 8, nrec: int = 100,
            seed: int = 5, square: float = 400.0, contrast: float = 0.05):
    """2-D adjoint workload: n x n quads of h, Stacey on all four sides;
    background (vp, vs, rho) = (3000, 1732, 2400), true model the background
    times (1 +- contrast) on a checkerboard of square-metre cells; nsrc Ricker
    forces (f0 12 Hz, vertical) at seeded points of the inner 80 %, nrec
    receivers on a line 100 m below the top.  Returns (true, start) workloads
    sharing sources and receivers."""
    L = n * h
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.1 * L, 0.9 * L, size=(nsrc, 2))
    src = [dict(x=float(x), z=float(z), angle=0.0, f0=12.0, factor=1e10, tshift=0.0) for x, z in pos]
    rec = [dict(x=float(x), z=L - 100.0) for x in np.linspace(0.05 * L, 0.95 * L, nrec)]
    ab = ("bottom", "right", "left", "top")
    cells = int(round(square / h))
    ez, ex = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    checker = (((ex // cells) + (ez // cells)) % 2).astype(np.int32)
    bg = (2400.0, 3000.0, 1732.0)
    true_mats = [elastic(*(x * (1 + contrast) for x in bg)), elastic(*(x * (1 - contrast) for x in bg))]
    true = Workload2D(f"2d_adjoint_{n}x{n}_true", quad_mesh(n, n, 0.0, L, 0.0, L, checker, true_mats, ab), src, rec,
                      dt, nstep, dt, {"courant": courant(dt, h, 3000.0 * (1 + contrast))})
    start = Workload2D(f"2d_adjoint_{n}x{n}", quad_mesh(n, n, 0.0, L, 0.0, L, None, [elastic(*bg)], ab), src, rec,
                       dt, nstep, dt, {"courant": courant(dt, h, 3000.0)})
    return true, start


@dataclass
class Workload3D:
    name: str
    mesh: dict
    sources: list
    receivers: list
    dt: float
    nstep: int
    dt_stated: float = 0.0
    notes: dict = field(default_factory=dict)


def hex_mesh(nx, ny, nz, h, material, rho, vp, vs, absorbing=(1, 1, 1, 1, 1, 0)):
    return dict(nx=nx, ny=ny, nz=nz, x0=0.0, y0=0.0, z0=0.0, hx=h, hy=h, hz=h,
                material=np.ascontiguousarray(material, dtype=np.int32), rho=np.asarray(rho, dtype=np.float64),
                vp=np.asarray(vp, dtype=np.float64), vs=np.asarray(vs, dtype=np.float64), absorbing=list(absorbing))


def moment_tensor(seed: int = 3):
    """Seeded random symmetric moment tensor of unit Frobenius norm."""
    rng = np.random.default_rng(seed)
    A = rng.standard_normal((3, 3))
    M = 0.5 * (A + A.T)
    return M / np.linalg.norm(M)


def config3(n: int = 64, nstep: int = 1000, h: float = 100.0, nrec_side: int = 20) -> Workload3D:
    """3-D layered elastic cube: n^3 hexahedra of h, 4 horizontal layers of
    n h / 4 from the top (vp 2000/3000/4500/6000 m/s, vs = vp / 1.73, rho
    2000/2300/2600/2900), free top, Stacey on the other five faces, seeded
    unit-norm moment tensor (Ricker 5 Hz) at 3 km depth under the centre,
    nrec_side^2 surface receivers.  The full configuration is n = 64."""
    L = n * h
    vp = np.array([2000.0, 3000.0, 4500.0, 6000.0])
    vs = vp / 1.73
    rho = np.array([2000.0, 2300.0, 2600.0, 2900.0])
    zc = (np.arange(n) + 0.5) * h  # element centre heights
    layer = np.minimum(3, ((L - zc) / (L / 4)).astype(np.int32))
    material = np.broadcast_to(layer[:, None, None], (n, n, n)).copy()
    mesh = hex_mesh(n, n, n, h, material, rho, vp, vs)
    M = moment_tensor(3)
    depth = min(3000.0, 0.47 * L)
    src = [dict(type="moment_tensor", x=L / 2, y=L / 2, z=L - depth, mxx=M[0, 0], myy=M[1, 1], mzz=M[2, 2],
                mxy=M[0, 1], mxz=M[0, 2], myz=M[1, 2], f0=5.0, factor=1e15, tshift=0.0)]
    spacing = L / nrec_side
    rec = [dict(x=spacing / 2 + spacing * i, y=spacing / 2 + spacing * j, z=L)
           for j in range(nrec_side) for i in range(nrec_side)]
    dt_stated = 2e-3
    dt = dt_stated if courant(dt_stated, h, vp.max()) <= 0.5 else 1e-3
    return Workload3D(f"3d_layered_{n}^3", mesh, src, rec, dt, nstep, dt_stated,
                      {"courant": courant(dt, h, vp.max()), "courant_stated": courant(dt_stated, h, vp.max())})


def nglob_3d(n: int) -> int:
    return (4 * n + 1) ** 3
