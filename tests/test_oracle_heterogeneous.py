"""Oracle pinning on the heterogeneous 2-D workload (CPU).

* Material placement: every GLL point of element e carries the density,
  bulk and shear moduli of material kmato[e] (one material per element,
  kappa = rho (vp^2 - 4/3 vs^2), mu = rho vs^2), and its mass weight
  w_i w_j J equals (h / 2)^2 w_i w_j on the regular mesh.
* Discrete reciprocity: mass, Stacey and stiffness matrices are symmetric
  and a force source and a displacement receiver at one point use the same
  Lagrange weights, so every step of the Newmark recursion has a symmetric
  transfer matrix and G(r <- s) = G(s <- r) component by component, with
  any medium and Stacey on every side.  A transposed operand, a wrong
  index or sign in the stiffness or boundary terms breaks it.  Checked in
  FP64 to 1e-10 relative L2 on both the zz and the xz/zx components
  (angle 0 is a force along -z, angle 90 along +x).
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from specfempp_amd import meshes


def _small():
    # 24 x 24 elements of 2.5 m (the full workload's spacing), field
    # correlation length 15 m so the medium varies between neighbours
    return meshes.config2(n=24, L=60.0, nstep=2500, nrec=4, seed=7, corr=15.0)


def test_material_placement_and_mass_weights():
    w = _small()
    o = oracle.Oracle2D(w.mesh, 8, 1, w.dt)
    rho, kappa, mu, wj = o.properties()
    d, cp, cs = meshes.material_arrays(w.mesh.materials)
    m = w.mesh.kmato
    assert len(np.unique(d[m])) == w.mesh.nspec  # one material per element
    np.testing.assert_allclose(rho, np.broadcast_to(d[m][:, None, None], rho.shape), rtol=1e-14)
    np.testing.assert_allclose(mu, np.broadcast_to((d * cs * cs)[m][:, None, None], mu.shape), rtol=1e-12)
    k = d * (cp * cp - 4.0 / 3.0 * cs * cs)
    np.testing.assert_allclose(kappa, np.broadcast_to(k[m][:, None, None], kappa.shape), rtol=1e-12)
    _, wq, _ = o.quadrature()
    h = w.notes["h"]
    np.testing.assert_allclose(wj, np.broadcast_to((h / 2) ** 2 * np.outer(wq, wq), wj.shape), rtol=1e-12)


@pytest.mark.parametrize("a,b", [((13.3, 21.7), (41.1, 37.9)), ((5.6, 52.2), (30.0, 30.0))])
def test_reciprocity_fp64(a, b):
    w = _small()
    o = oracle.Oracle2D(w.mesh, 8, 1, w.dt)
    pa, pb = dict(x=a[0], z=a[1]), dict(x=b[0], z=b[1])

    def run(at, angle, rec):
        src = [dict(x=at["x"], z=at["z"], angle=angle, f0=15.0, factor=1e10, tshift=0.0)]
        s, _ = o.forward(src, rec, w.nstep, w.dt, 1, ("displacement",))
        return s[:, 0]  # (nsig, nrec, 2)

    z_at_a = run(pa, 0.0, [pb])
    z_at_b = run(pb, 0.0, [pa])
    x_at_a = run(pa, 90.0, [pb])
    assert np.linalg.norm(z_at_a[:, 0, 1]) > 0
    assert rel_l2(z_at_b[:, 0, 1], z_at_a[:, 0, 1]) <= 1e-10  # G_zz(b, a) = G_zz(a, b)
    assert rel_l2(-z_at_b[:, 0, 0], x_at_a[:, 0, 1]) <= 1e-10  # G_zx(b, a) = G_xz(a, b)
