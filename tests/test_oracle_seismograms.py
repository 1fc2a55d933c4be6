"""Oracle pinning of seismogram sampling (CPU).

On a linear displacement field u = (a x + b z, c x + d z) the GLL derivatives
and the Lagrange interpolation are exact, so at every receiver the
displacement seismogram is u(x_r, z_r) and the pressure seismogram
(wavefield.hpp:49-85) is -(lambda + 2/3 mu)(a + d) = -kappa (a + d), with a zero
second component.  Checked on the reference's 80 x 60 mesh database.
"""
import numpy as np

import oracle
from conftest import load_case


def _linear_field(o, coef):
    a, b, c, d = coef
    coord, _ = o.geometry()
    ibool, midx, _ = o.numbering()
    xg = np.zeros(o.nglob)
    zg = np.zeros(o.nglob)
    g = midx[ibool.ravel()]
    xg[g] = coord[..., 0].ravel()
    zg[g] = coord[..., 1].ravel()
    fields = np.zeros((3, 2, o.nglob))
    fields[0, 0] = a * xg + b * zg
    fields[0, 1] = c * xg + d * zg
    return fields


def test_pressure_and_displacement_of_linear_field():
    mesh, cfg = load_case("elastic_free_80x60")
    o = oracle.Oracle2D(mesh, 8, 1, cfg["dt"])
    coef = (2e-4, -3e-4, 5e-4, 7e-4)
    fields = _linear_field(o, coef)
    rec = cfg["stations"][:10] + [dict(x=1234.5, z=987.6), dict(x=3001.0, z=1700.25)]
    out = o.sample(rec, fields, ("displacement", "pressure"))
    m = mesh.materials[0]
    kappa = m.density * (m.cp ** 2 - 4.0 / 3.0 * m.cs ** 2)
    a, b, c, d = coef
    for r, pt in enumerate(rec):
        ux, uz = a * pt["x"] + b * pt["z"], c * pt["x"] + d * pt["z"]
        assert abs(out[0, r, 0] - ux) <= 1e-9 * max(abs(ux), 1.0)
        assert abs(out[0, r, 1] - uz) <= 1e-9 * max(abs(uz), 1.0)
        assert abs(out[1, r, 0] - (-kappa * (a + d))) <= 1e-9 * abs(kappa * (a + d))
        assert out[1, r, 1] == 0.0
