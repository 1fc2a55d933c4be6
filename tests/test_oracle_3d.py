"""Pins of the 3-D oracle's time loop and of the moment-tensor sources.

- The threaded element loop gives the same bits as the serial one.
- Without absorbing faces the central-difference (Newmark beta = 0,
  gamma = 1/2) energy E = 1/2 v+^T M v+ + 1/2 u_{n+1}^T K u_n, with
  v+ = v_n + dt/2 a_n and u_{n+1} = u_n + dt v+, is conserved once the source
  has died out.  A wrong Newmark factor, mass or a non-symmetric stiffness
  breaks it.
- A moment tensor is the limit of force dipoles: M_cr d/dx_r h(x_s) equals
  (h(x_s + d/2 e_r) - h(x_s - d/2 e_r)) M_cr / d up to O(d^2), and every
  seismogram is linear in the source array.  Checked in 3-D and in 2-D (the
  2-D moment-tensor source, moment_tensor_source.cpp:88-130, has no fixture
  in the reference's tests).
"""
import math

import numpy as np
import pytest

import oracle
from specfempp_amd import meshes


def small_cube(n=4, h=100.0, absorbing=(1, 1, 1, 1, 1, 0)):
    mat = np.zeros((n, n, n), dtype=np.int32)
    return meshes.hex_mesh(n, n, n, h, mat, [2500.0], [3000.0], [1730.0], absorbing=absorbing)


@pytest.mark.parametrize("precision", [4, 8])
def test_threaded_forward_is_bitwise_serial(precision):
    w = meshes.config3(n=4, nstep=40, h=100.0, nrec_side=2)
    o = oracle.Oracle3D(w.mesh, precision, w.dt)
    types = ("displacement", "velocity", "acceleration")
    s1, _, f1 = o.forward(w.sources, w.receivers, w.nstep, w.dt, 1, types, return_fields=True, nthreads=1)
    s4, _, f4 = o.forward(w.sources, w.receivers, w.nstep, w.dt, 1, types, return_fields=True, nthreads=4)
    assert np.abs(f1).max() > 0
    assert np.array_equal(s1, s4) and np.array_equal(f1, f4)


def test_energy_is_conserved_without_absorbing_faces():
    n, h, dt = 3, 100.0, 1e-3
    mesh = small_cube(n, h, absorbing=(0, 0, 0, 0, 0, 0))
    o = oracle.Oracle3D(mesh, 8, dt)
    src = [dict(type="force", x=140.0, y=160.0, z=130.0, fx=0.3, fy=-0.5, fz=1.0, f0=25.0, factor=1e9)]
    rec = [dict(x=50.0, y=50.0, z=50.0)]
    minv = o.mass_inverse()

    def energy(nstep):
        _, _, f = o.forward(src, rec, nstep, dt, 1, return_fields=True)
        u, v, a = f
        vh = v + 0.5 * dt * a
        u1 = u + dt * vh
        Ku = -o.apply_stiffness(u)  # apply_stiffness gives -K u
        kin = 0.5 * np.sum(vh * vh / minv)
        pot = 0.5 * np.sum(u1 * Ku)
        return kin + pot, kin

    # the 25 Hz Ricker starts at t0 = -1.2 / f0 and is below 1e-16 of its peak after 0.15 s
    e1, k1 = energy(200)
    e2, k2 = energy(330)
    assert k1 > 0 and abs(k1 - k2) > 1e-3 * k1  # energy moves between kinetic and potential
    assert abs(e2 - e1) <= 1e-10 * abs(e1)


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def test_moment_tensor_is_the_force_dipole_limit_3d():
    mesh = small_cube(3, 100.0)
    o = oracle.Oracle3D(mesh, 8, 1e-3)
    M = meshes.moment_tensor(7)
    xs = np.array([140.0, 160.0, 130.0])  # inside element (1, 1, 1), off its GLL lines
    rec = [dict(x=40.0, y=60.0, z=290.0), dict(x=260.0, y=220.0, z=100.0)]
    base = dict(f0=12.0, tshift=0.0)
    mt = [dict(type="moment_tensor", x=xs[0], y=xs[1], z=xs[2], mxx=M[0, 0], myy=M[1, 1], mzz=M[2, 2],
               mxy=M[0, 1], mxz=M[0, 2], myz=M[1, 2], factor=1.0, **base)]
    d = 1e-4 * 100.0
    dip = []
    for c in range(3):
        for r in range(3):
            for sgn in (1.0, -1.0):
                p = xs.copy()
                p[r] += sgn * d / 2
                f = [0.0, 0.0, 0.0]
                f[c] = sgn * M[c, r] / d
                dip.append(dict(type="force", x=p[0], y=p[1], z=p[2], fx=f[0], fy=f[1], fz=f[2], factor=1.0, **base))
    s_mt, _ = o.forward(mt, rec, 300, 1e-3)
    s_dip, _ = o.forward(dip, rec, 300, 1e-3)
    assert np.abs(s_mt).max() > 0
    assert _rel(s_dip, s_mt) < 1e-6
    # the transposed tensor of a non-symmetric test would differ: a wrong index fails
    bad = [dict(mt[0], mxy=-M[0, 1])]
    assert _rel(o.forward(bad, rec, 300, 1e-3)[0], s_mt) > 1e-2


def test_moment_tensor_is_the_force_dipole_limit_2d():
    n, h = 6, 50.0
    mesh = meshes.quad_mesh(n, n, 0.0, n * h, 0.0, n * h, materials=[meshes.elastic(2200.0, 2500.0, 1443.0)],
                            absorbing=("bottom", "right", "left"))
    o = oracle.Oracle2D(mesh, 8)
    Mxx, Mzz, Mxz = 0.7, -0.4, 0.55
    xs, zs = 137.0, 161.0
    base = dict(f0=15.0, tshift=0.0)
    mt = [dict(type="moment_tensor", x=xs, z=zs, mxx=Mxx, mzz=Mzz, mxz=Mxz, factor=1.0, **base)]
    rec = [dict(x=40.0, z=280.0), dict(x=260.0, z=60.0)]
    d = 1e-4 * h
    # force direction (sin angle, -cos angle): +x is angle 90, +z is angle 180
    ang = {0: 90.0, 1: 180.0}
    Mm = [[Mxx, Mxz], [Mxz, Mzz]]
    dip = []
    for c in range(2):
        for r in range(2):
            for sgn in (1.0, -1.0):
                p = [xs, zs]
                p[r] += sgn * d / 2
                dip.append(dict(x=p[0], z=p[1], angle=ang[c], factor=sgn * Mm[c][r] / d, **base))
    s_mt, _ = o.forward(mt, rec, 400, 2e-4)
    s_dip, _ = o.forward(dip, rec, 400, 2e-4)
    assert np.abs(s_mt).max() > 0
    assert _rel(s_dip, s_mt) < 1e-6
    bad = [dict(mt[0], mxz=-Mxz)]
    assert _rel(o.forward(bad, rec, 400, 2e-4)[0], s_mt) > 1e-2


W5 = np.array([1 / 10, 49 / 90, 32 / 45, 49 / 90, 1 / 10])  # 5-point GLL weights


def _lattice(o, n):
    """Lattice coordinates (I, J, K) of every global point of an n^3 cube."""
    ib = o.ibool()  # (nspec, k, j, i)
    I = np.zeros(o.nglob, dtype=np.int64)
    J, K = I.copy(), I.copy()
    e = np.arange(n ** 3)
    ex, ey, ez = e % n, (e // n) % n, e // (n * n)
    loc = np.arange(5)
    I[ib] = 4 * ex[:, None, None, None] + loc[None, None, None, :]
    J[ib] = 4 * ey[:, None, None, None] + loc[None, None, :, None]
    K[ib] = 4 * ez[:, None, None, None] + loc[None, :, None, None]
    return I, J, K


def _assembled_weights(n):
    W = np.zeros(4 * n + 1)
    for e in range(n):
        W[4 * e:4 * e + 5] += W5
    return W


def _stacey_damping_closed_form(o, n, h, rho, vp, vs):
    """Per global point and component, the Stacey damping C on an all-absorbing
    n^3 box: on each face through the point, rho * (vp normal, vs tangential)
    times the in-face GLL weights and the face jacobian (h / 2)^2
    (stacey.tpp:36-72 on axis-aligned faces)."""
    I, J, K = _lattice(o, n)
    W = _assembled_weights(n)
    L = 4 * n
    C = np.zeros((3, o.nglob))
    coords = (I, J, K)
    for axis in range(3):
        others = [a for a in range(3) if a != axis]
        wface = W[coords[others[0]]] * W[coords[others[1]]] * (h / 2) ** 2
        for side in (0, L):
            on = coords[axis] == side
            for c in range(3):
                C[c, on] += rho * (vp if c == axis else vs) * wface[on]
    return C


def test_stacey_mass_term_matches_the_closed_form():
    n, h, dt = 3, 100.0, 1e-3
    rho, vp, vs = 2500.0, 3000.0, 1730.0
    o_st = oracle.Oracle3D(small_cube(n, h, absorbing=(1,) * 6), 8, dt)
    o_fr = oracle.Oracle3D(small_cube(n, h, absorbing=(0,) * 6), 8, dt)
    C_oracle = (1 / o_st.mass_inverse() - 1 / o_fr.mass_inverse()) * 2 / dt
    C = _stacey_damping_closed_form(o_st, n, h, rho, vp, vs)
    assert np.count_nonzero(C) > 0
    assert np.abs(C_oracle - C).max() <= 1e-9 * np.abs(C).max()
    # and the free mass is the closed form rho w_i w_j w_k (h / 2)^3, assembled
    I, J, K = _lattice(o_fr, n)
    W = _assembled_weights(n)
    M = rho * W[I] * W[J] * W[K] * (h / 2) ** 3
    assert np.abs(1 / o_fr.mass_inverse() - M[None]).max() <= 1e-9 * M.max()


def test_stacey_energy_balance():
    """Source-free steps with Stacey on all faces: the mass term makes the
    damping act on the mean of v(n+1/2) and v(n+3/2), so
    E(n+3/2) - E(n+1/2) = -dt/4 (v(n+3/2) + v(n+1/2))^T C (v(n+3/2) + v(n+1/2))
    with E as in the free case and C the closed-form damping."""
    n, h, dt = 3, 100.0, 1e-3
    rho, vp, vs = 2500.0, 3000.0, 1730.0
    o = oracle.Oracle3D(small_cube(n, h, absorbing=(1,) * 6), 8, dt)
    o_fr = oracle.Oracle3D(small_cube(n, h, absorbing=(0,) * 6), 8, dt)
    M = 1 / o_fr.mass_inverse()
    C = _stacey_damping_closed_form(o, n, h, rho, vp, vs)
    src = [dict(type="force", x=140.0, y=160.0, z=130.0, fx=0.3, fy=-0.5, fz=1.0, f0=25.0, factor=1e9)]
    rec = [dict(x=50.0, y=50.0, z=50.0)]

    def state(nstep):
        _, _, (u, v, a) = o.forward(src, rec, nstep, dt, 1, return_fields=True)
        vh = v + 0.5 * dt * a
        u1 = u + dt * vh
        E = 0.5 * np.sum(vh * M * vh) + 0.5 * np.sum(u1 * -o.apply_stiffness(u))
        return E, vh

    for N in (170, 230):
        E0, vh0 = state(N)
        E1, vh1 = state(N + 1)
        s = vh0 + vh1
        pred = -dt / 4 * np.sum(s * C * s)
        assert pred < 0 and abs(pred) > 1e-5 * E0
        assert abs((E1 - E0) - pred) <= 1e-8 * abs(pred)
