"""Pin the oracle's stiffness operator against an independent dense weak-form
assembly written here in numpy: K_e = sum_q W_q B_q^T D B_q with Lagrange
basis gradients (2-D plane strain and 3-D), plus symmetry and rigid-body
null space.  A dropped term, a wrong sign or a transposed derivative matrix in
the oracle fails these."""
import numpy as np
import pytest

import oracle
from specfempp_amd import meshes

XI = np.array([-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0])
W = np.array([0.1, 49 / 90, 32 / 45, 49 / 90, 0.1])


def lagrange_deriv_matrix():
    # D[i, j] = l_j'(xi_i) by differentiating the monomial interpolant
    V = np.vander(XI, 5, increasing=True)
    C = np.linalg.inv(V)  # columns: coefficients of l_j
    dV = np.zeros_like(V)
    for k in range(1, 5):
        dV[:, k] = k * XI ** (k - 1)
    return dV @ C


def iso_D(dim, lam, mu):
    if dim == 2:
        return np.array([[lam + 2 * mu, lam, 0], [lam, lam + 2 * mu, 0], [0, 0, mu]])
    D = np.zeros((6, 6))
    D[:3, :3] = lam
    D[np.arange(3), np.arange(3)] = lam + 2 * mu
    D[3, 3] = D[4, 4] = D[5, 5] = mu
    return D


def dense_K_2d(nx, nz, h, rho, vp, vs):
    """Global dense K of a structured quad mesh in the oracle's numbering."""
    mesh = meshes.quad_mesh(nx, nz, 0, nx * h, 0, nz * h, materials=[meshes.elastic(rho, vp, vs)], absorbing=())
    o = oracle.Oracle2D(mesh, 8)
    ibool, midx, c2m = o.numbering()
    _, (xix, xiz, gammax, gammaz, jac) = o.geometry()
    Dm = lagrange_deriv_matrix()
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    D = iso_D(2, lam, mu)
    n = o.nglob
    K = np.zeros((2 * n, 2 * n))
    for e in range(mesh.nspec):
        dofs = np.concatenate([midx[ibool[e].ravel()], n + midx[ibool[e].ravel()]])
        Ke = np.zeros((50, 50))
        for iz in range(5):
            for ix in range(5):
                # d phi_a / d xi, d gamma for a = (bz, bx)
                dphi_dxi = np.zeros(25)
                dphi_dga = np.zeros(25)
                for bz in range(5):
                    for bx in range(5):
                        a = bz * 5 + bx
                        dphi_dxi[a] = Dm[ix, bx] * (bz == iz)
                        dphi_dga[a] = Dm[iz, bz] * (bx == ix)
                dx = xix[e, iz, ix] * dphi_dxi + gammax[e, iz, ix] * dphi_dga
                dz = xiz[e, iz, ix] * dphi_dxi + gammaz[e, iz, ix] * dphi_dga
                B = np.zeros((3, 50))
                B[0, :25] = dx
                B[1, 25:] = dz
                B[2, :25] = dz
                B[2, 25:] = dx
                Ke += W[ix] * W[iz] * jac[e, iz, ix] * B.T @ D @ B
        K[np.ix_(dofs, dofs)] += Ke
    return o, K


@pytest.mark.parametrize("nx,nz", [(1, 1), (2, 3)])
def test_oracle_2d_stiffness_matches_dense_weak_form(nx, nz):
    o, K = dense_K_2d(nx, nz, 100.0, 2000.0, 3000.0, 1700.0)
    rng = np.random.default_rng(7)
    u = rng.standard_normal((2, o.nglob))
    ours = o.apply_stiffness(u, np.zeros_like(u))
    ref = -(K @ u.ravel()).reshape(2, -1)
    assert np.abs(ours - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())


def test_oracle_2d_rigid_body_modes():
    o, K = dense_K_2d(2, 2, 50.0, 2200.0, 2500.0, 1443.0)
    _, midx, _ = o.numbering()
    coord, _ = o.geometry()
    ibool, _, _ = o.numbering()
    xy = np.zeros((o.nglob, 2))
    xy[midx[ibool.ravel()]] = coord.reshape(-1, 2)
    scale = np.abs(o.apply_stiffness(np.stack([xy[:, 0], np.zeros(o.nglob)]), np.zeros((2, o.nglob)))).max()
    for u in (np.stack([np.ones(o.nglob), np.zeros(o.nglob)]), np.stack([-xy[:, 1], xy[:, 0]])):
        r = o.apply_stiffness(u, np.zeros_like(u))
        assert np.abs(r).max() < 1e-9 * scale


def dense_K_3d(nx, ny, nz, h, rho, vp, vs):
    mesh = meshes.hex_mesh(nx, ny, nz, h, np.zeros((nz, ny, nx)), [rho], [vp], [vs], absorbing=(0,) * 6)
    o = oracle.Oracle3D(mesh, 8)
    ib = o.ibool()
    Dm = lagrange_deriv_matrix()
    mu = rho * vs * vs
    lam = rho * vp * vp - 2 * mu
    D = iso_D(3, lam, mu)
    n = o.nglob
    K = np.zeros((3 * n, 3 * n))
    J = h ** 3 / 8.0
    g = 2.0 / h
    for e in range(nx * ny * nz):
        loc = ib[e].ravel()  # p = (k, j, i)
        dofs = np.concatenate([loc, n + loc, 2 * n + loc])
        Ke = np.zeros((375, 375))
        for k in range(5):
            for j in range(5):
                for i in range(5):
                    d = np.zeros((3, 125))
                    for c in range(5):
                        d[0, (k * 5 + j) * 5 + c] = Dm[i, c] * g
                        d[1, (k * 5 + c) * 5 + i] = Dm[j, c] * g
                        d[2, (c * 5 + j) * 5 + i] = Dm[k, c] * g
                    B = np.zeros((6, 375))
                    for r in range(3):
                        B[r, r * 125:(r + 1) * 125] = d[r]
                    B[3, 0:125], B[3, 125:250] = d[1], d[0]          # xy
                    B[4, 0:125], B[4, 250:375] = d[2], d[0]          # xz
                    B[5, 125:250], B[5, 250:375] = d[2], d[1]        # yz
                    Ke += W[i] * W[j] * W[k] * J * B.T @ D @ B
        K[np.ix_(dofs, dofs)] += Ke
    return o, K


def test_oracle_3d_stiffness_matches_dense_weak_form():
    o, K = dense_K_3d(2, 1, 1, 100.0, 2600.0, 4500.0, 2600.0)
    rng = np.random.default_rng(11)
    u = rng.standard_normal((3, o.nglob))
    ours = o.apply_stiffness(u)
    ref = -(K @ u.ravel()).reshape(3, -1)
    assert np.abs(ours - ref).max() <= 1e-10 * np.abs(ref).max()
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())
    # rigid translations and rotations are in the null space
    lat = np.zeros((o.nglob, 3))
    ib = o.ibool()
    pts = np.array([[(4 * (e % 2) + i), j + 0.0, k + 0.0] for e in range(2) for k in range(5) for j in range(5)
                    for i in range(5)])
    lat[ib.ravel()] = pts
    scale = np.abs(K).max()
    for c in range(3):
        t = np.zeros((3, o.nglob))
        t[c] = 1.0
        assert np.abs(K @ t.ravel()).max() < 1e-10 * scale * o.nglob
