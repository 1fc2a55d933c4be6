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


# GLL points of NGLL = 5 on [-1, 1] and their weights (closed form)
_GLL_X = np.array([-1.0, -np.sqrt(3.0 / 7.0), 0.0, np.sqrt(3.0 / 7.0), 1.0])
_GLL_W = np.array([0.1, 49.0 / 90.0, 32.0 / 45.0, 49.0 / 90.0, 0.1])


def hex_database(nx, ny, nz, h, material, rho, vp, vs, absorbing=(1, 1, 1, 1, 1, 0), perturb=0.0, seed=0,
                 dt=1e-3):
    """A 3-D mesh database (specfempp_amd.io.database3d.Mesh3DDatabase) in
    the layout generate_databases writes and the reference reads
    (src/io/mesh/impl/fortran/dim3/mesh.cpp): 8-node hexahedra on an nx x ny
    x nz grid of spacing h whose interior corner nodes in the lower half along
    x (node index 1 .. nx // 2) are moved by up to perturb * h along each axis
    (seeded), so the elements touching them are irregular; the others are
    cubes, described by xix_regular and jacobian_regular.

    Per GLL point: the trilinear map's position, derivatives (xix ... gammaz)
    and jacobian, evaluated in double; kappa, mu, rho, rho vp, rho vs of the
    element's material (vs > 0).  Absorbing faces (xmin, xmax, ymin, ymax,
    bottom, top flags): every face point's index, unit outward normal and
    jacobian2Dw (surface jacobian times the two in-face GLL weights); the top
    faces that are not absorbing form the free surface.  Mass matrix (rho w^3
    J assembled) and the Stacey mass terms dt/2 jacobian2Dw (rho vp n_c^2 +
    rho vs (1 - n_c^2)); single-process inner/outer arrays (every element
    inner), surface flags and element adjacency (elements sharing a point).
    Numbering: first visit in (element, k, j, i) order, elements x fastest."""
    from .io.database3d import Faces, Mesh3DDatabase

    material = np.asarray(material, dtype=np.int32).reshape(nz, ny, nx)
    rho, vp, vs = (np.asarray(v, dtype=np.float64) for v in (rho, vp, vs))
    if np.any(vs <= 0):
        raise ValueError("hex_database builds elastic meshes (vs > 0)")
    rng = np.random.default_rng(seed)
    # corner nodes, interior ones perturbed
    X = np.zeros((nz + 1, ny + 1, nx + 1, 3))
    X[..., 0] = (np.arange(nx + 1) * h)[None, None, :]
    X[..., 1] = (np.arange(ny + 1) * h)[None, :, None]
    X[..., 2] = (np.arange(nz + 1) * h)[:, None, None]
    moved = np.zeros((nz + 1, ny + 1, nx + 1), dtype=bool)
    if perturb > 0:
        inner = (slice(1, nz), slice(1, ny), slice(1, max(2, nx // 2 + 1)))
        X[inner] += rng.uniform(-perturb * h, perturb * h, size=X[inner].shape)
        moved[inner] = True
    nspec = nx * ny * nz
    # numbering: first visit on the GLL lattice
    LX, LY, LZ = 4 * nx + 1, 4 * ny + 1, 4 * nz + 1
    first = -np.ones(LX * LY * LZ, dtype=np.int64)
    ibool = np.zeros((nspec, 5, 5, 5), dtype=np.int32)
    inum = 0
    for e in range(nspec):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        for k in range(5):
            for j in range(5):
                for i in range(5):
                    L = ((4 * ez + k) * LY + (4 * ey + j)) * LX + (4 * ex + i)
                    if first[L] < 0:
                        first[L] = inum
                        inum += 1
                    ibool[e, k, j, i] = first[L]
    nglob = inum
    xi = _GLL_X
    # trilinear shape functions and derivatives at the 125 points: (8, 125), (3, 8, 125)
    sgn = np.array([[(c & 1) * 2 - 1, ((c >> 1) & 1) * 2 - 1, ((c >> 2) & 1) * 2 - 1] for c in range(8)], float)
    K, J_, I = np.meshgrid(np.arange(5), np.arange(5), np.arange(5), indexing="ij")
    r = np.stack([xi[I.ravel()], xi[J_.ravel()], xi[K.ravel()]])  # (3, 125)
    f = 1 + sgn[:, :, None] * r[None, :, :]                      # (8, 3, 125)
    Nsh = f[:, 0] * f[:, 1] * f[:, 2] / 8
    dN = np.stack([sgn[:, 0, None] * f[:, 1] * f[:, 2], f[:, 0] * sgn[:, 1, None] * f[:, 2],
                   f[:, 0] * f[:, 1] * sgn[:, 2, None]]) / 8         # (3, 8, 125)
    pts = np.zeros((nglob, 3))
    derivs = {n: np.zeros((nspec, 5, 5, 5), np.float32) for n in
              ("xix", "xiy", "xiz", "etax", "etay", "etaz", "gammax", "gammay", "gammaz", "jacobian")}
    jac64 = np.zeros((nspec, 125))
    inv64 = np.zeros((nspec, 125, 3, 3))  # [s][r] = d xi_s / d x_r
    regular = np.zeros(nspec, dtype=bool)
    for e in range(nspec):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        C = np.array([X[ez + ((c >> 2) & 1), ey + ((c >> 1) & 1), ex + (c & 1)] for c in range(8)])  # (8, 3)
        corner_moved = any(moved[ez + ((c >> 2) & 1), ey + ((c >> 1) & 1), ex + (c & 1)] for c in range(8))
        regular[e] = not corner_moved
        P = Nsh.T @ C                                   # (125, 3)
        pts[ibool[e].ravel()] = P
        dX = np.einsum("snp,nr->prs", dN, C)            # (125, 3 r, 3 s): d x_r / d xi_s
        Jd = np.linalg.det(dX)
        inv = np.linalg.inv(dX)                         # (125, 3 s, 3 r)
        jac64[e] = Jd
        inv64[e] = inv
        for s_, nm in enumerate(("xi", "eta", "gamma")):
            for r_, ax in enumerate("xyz"):
                derivs[nm + ax][e] = inv[:, s_, r_].reshape(5, 5, 5)
        derivs["jacobian"][e] = Jd.reshape(5, 5, 5)
    mat = material.ravel()
    rho_p = np.repeat(rho[mat], 125).reshape(nspec, 5, 5, 5).astype(np.float32)
    mu_p = np.repeat((rho * vs * vs)[mat], 125).reshape(nspec, 5, 5, 5).astype(np.float32)
    kap_p = np.repeat((rho * (vp * vp - 4.0 / 3.0 * vs * vs))[mat], 125).reshape(nspec, 5, 5, 5).astype(np.float32)
    rvp = np.repeat((rho * vp)[mat], 125).reshape(nspec, 5, 5, 5).astype(np.float32)
    rvs = np.repeat((rho * vs)[mat], 125).reshape(nspec, 5, 5, 5).astype(np.float32)
    w3 = (_GLL_W[I] * _GLL_W[J_] * _GLL_W[K]).ravel()
    mass = np.zeros(nglob)
    np.add.at(mass, ibool.reshape(nspec, 125), rho[mat][:, None] * w3[None, :] * jac64)

    # faces: (side, element test, fixed axis, fixed index, outward sign)
    sides = [("xmin", 0, 0, -1), ("xmax", 0, 4, 1), ("ymin", 1, 0, -1), ("ymax", 1, 4, 1), ("bottom", 2, 0, -1),
             ("top", 2, 4, 1)]

    def face_points(e, axis, fixed, sign):
        ijk = np.zeros((3, 25), dtype=np.int32)
        nrm = np.zeros((3, 25))
        jw = np.zeros(25)
        o1, o2 = [a for a in range(3) if a != axis]
        q = 0
        for b in range(5):
            for a in range(5):
                idx = [0, 0, 0]
                idx[axis], idx[o1], idx[o2] = fixed, a, b
                i_, j2, k_ = idx
                p = (k_ * 5 + j2) * 5 + i_
                # Nanson: the area vector is J grad(xi_axis) dA_ref
                n = sign * jac64[e, p] * inv64[e, p, axis, :]
                area = np.linalg.norm(n)
                ijk[:, q] = idx
                nrm[:, q] = n / area
                jw[q] = area * _GLL_W[a] * _GLL_W[b]
                q += 1
        return ijk, nrm, jw

    abs_f, free_f, ibelm = [], [], {}
    for side, axis, fixed, sign in sides:
        flag = absorbing[["xmin", "xmax", "ymin", "ymax", "bottom", "top"].index(side)]
        lst = []
        for e in range(nspec):
            ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
            pos = (ex, ey, ez)[axis]
            last = (nx, ny, nz)[axis] - 1
            if pos != (0 if fixed == 0 else last):
                continue
            lst.append(e)
            if flag:
                abs_f.append((e,) + face_points(e, axis, fixed, sign))
            elif side == "top":
                free_f.append((e,) + face_points(e, axis, fixed, sign))
        ibelm[side] = np.array(lst if flag else [], dtype=np.int32)

    def faces(lst):
        if not lst:
            return Faces.empty()
        return Faces(np.array([f_[0] for f_ in lst], np.int32), np.array([f_[1] for f_ in lst], np.int32),
                     np.array([f_[3] for f_ in lst], np.float32), np.array([f_[2] for f_ in lst], np.float32))

    absorbing_faces, free_surface = faces(abs_f), faces(free_f)
    abs_mass = np.zeros((3, nglob))
    for e, ijk, nrm, jw in abs_f:
        for q in range(25):
            i_, j2, k_ = ijk[:, q]
            g = ibool[e, k_, j2, i_]
            m = mat[e]
            for c in range(3):
                abs_mass[c, g] += dt / 2 * jw[q] * (rho[m] * vp[m] * nrm[c, q] ** 2 +
                                                    rho[m] * vs[m] * (1 - nrm[c, q] ** 2))
    # adjacency: elements sharing a point
    owners = [[] for _ in range(nglob)]
    for e in range(nspec):
        for g in np.unique(ibool[e]):
            owners[g].append(e)
    xadj, adj = [0], []
    for e in range(nspec):
        nb = sorted({o for g in np.unique(ibool[e]) for o in owners[g]} - {e})
        adj += nb
        xadj.append(len(adj))
    surface_elems = np.zeros(nspec, dtype=bool)
    for lst in ibelm.values():
        surface_elems[lst] = True
    for e in range(nspec):
        ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
        surface_elems[e] = ex in (0, nx - 1) or ey in (0, ny - 1) or ez in (0, nz - 1)
    lat = np.zeros(nglob, dtype=bool)
    for e in range(nspec):
        for k in range(5):
            for j in range(5):
                for i in range(5):
                    ex, ey, ez = e % nx, (e // nx) % ny, e // (nx * ny)
                    I0, J0, K0 = 4 * ex + i, 4 * ey + j, 4 * ez + k
                    if I0 in (0, LX - 1) or J0 in (0, LY - 1) or K0 in (0, LZ - 1):
                        lat[ibool[e, k, j, i]] = True
    nsurf_faces = 2 * (nx * ny + ny * nz + nx * nz)
    irregular = np.flatnonzero(~regular).astype(np.int32)
    params = dict(acoustic_simulation=False, elastic_simulation=True, poroelastic_simulation=False,
                  anisotropy=False, stacey_abc=bool(any(absorbing)), pml_abc=False, approximate_ocean_load=False,
                  use_mesh_coloring=False, ndim=3, ngllx=5, nglly=5, ngllz=5, ngllsquare=25, nproc=1,
                  nspec=nspec, nspec_poro=0, nglob=nglob, nglob_ocean=0,
                  nspec2D_bottom=len(ibelm["bottom"]), nspec2D_top=len(ibelm["top"]),
                  nspec2D_xmin=len(ibelm["xmin"]), nspec2D_xmax=len(ibelm["xmax"]),
                  nspec2D_ymin=len(ibelm["ymin"]), nspec2D_ymax=len(ibelm["ymax"]), nspec_irregular=len(irregular),
                  num_neighbors=len(adj), nfaces_surface=nsurf_faces,
                  num_abs_boundary_faces=len(absorbing_faces.ispec),
                  num_free_surface_faces=len(free_surface.ispec), num_coupling_ac_el_faces=0,
                  num_coupling_ac_po_faces=0, num_coupling_el_po_faces=0, num_coupling_po_el_faces=0,
                  num_interfaces_ext_mesh=0, max_nibool_interfaces_ext_mesh=0,
                  nspec_inner_acoustic=0, nspec_outer_acoustic=0, nspec_inner_elastic=nspec,
                  nspec_outer_elastic=0, nspec_inner_poroelastic=0, nspec_outer_poroelastic=0,
                  num_phase_ispec_acoustic=0, num_phase_ispec_elastic=nspec, num_phase_ispec_poroelastic=0,
                  num_colors_inner_acoustic=0, num_colors_outer_acoustic=0, num_colors_inner_elastic=0,
                  num_colors_outer_elastic=0)
    phase = np.zeros((nspec, 2), dtype=np.int32)
    phase[:, 1] = np.arange(nspec)
    db = Mesh3DDatabase(params, ibool, pts[:, 0].astype(np.float32), pts[:, 1].astype(np.float32),
                        pts[:, 2].astype(np.float32), irregular, np.float32(2.0 / h), np.float32((h / 2) ** 3),
                        derivs, kap_p, mu_p, rho_p, np.ones(nspec, dtype=np.int32))
    db.mass_elastic = mass.astype(np.float32)
    db.rho_vp, db.rho_vs = rvp, rvs
    db.absorbing = absorbing_faces
    db.abs_mass_elastic = tuple(abs_mass[c].astype(np.float32) for c in range(3))
    db.ibelm = ibelm
    db.free_surface = free_surface
    db.ispec_is_inner = np.ones(nspec, dtype=bool)
    db.inner_outer = {"elastic": (nspec, 0, phase)}
    db.ispec_is_surface = surface_elems
    db.iglob_is_surface = lat
    db.neighbors_xadj = np.array(xadj, dtype=np.int32)
    db.neighbors_adjncy = np.array(adj, dtype=np.int32)
    return db
