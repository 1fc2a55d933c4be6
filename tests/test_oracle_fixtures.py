"""The oracle against the reference's own fixtures and known answers."""
import os

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, load_case, rel_l1_per_station

TYPES = {"d": "displacement", "v": "velocity", "a": "acceleration", "p": "pressure"}


def test_gll_known_answers():
    # tests/unit-tests/gll/gll_tests.cpp:165-178 (zwgljd, np = 5, alpha = beta = 0)
    mesh, _ = load_case("elastic_free_80x60")
    xi, w, hp = oracle.Oracle2D(mesh, 8).quadrature()
    np.testing.assert_allclose(xi, [-1.0, -0.6546536707, 0.0, 0.6546536707, 1.0], atol=1e-6)
    np.testing.assert_allclose(w, [0.1, 0.5444444444, 0.7111111111, 0.5444444444, 0.1], atol=1e-6)
    # tests/unit-tests/lagrange/Lagrange_tests.cpp:9-47: hprime(i, j) = l_j'(xi_i),
    # corner values -/+ N(N+1)/4 with N = 4, zero on the interior diagonal.
    assert hp[0, 0] == pytest.approx(-5.0)
    assert hp[4, 4] == pytest.approx(5.0)
    assert np.allclose(np.diag(hp)[1:4], 0.0)
    # the derivative of a constant vanishes; of a linear function is 1
    np.testing.assert_allclose(hp.sum(axis=1), 0.0, atol=1e-12)
    np.testing.assert_allclose(hp @ xi, 1.0, atol=1e-12)
    # GLL quadrature is exact for degree <= 2N-3 = 5
    for k in range(6):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert np.dot(w, xi ** k) == pytest.approx(exact, abs=1e-12)


def test_mass_inverse_matches_specfem2d():
    # tests/unit-tests/domain/rmass_inverse_tests.cpp:176-234 (rel. L1 <= 1e-5)
    mesh, cfg = load_case("mass_80x60")
    d = np.load(os.path.join(GOLDEN, "mass_80x60", "mass.npz"))
    ref_ibool = d["index_mapping"].reshape(mesh.nspec, 5, 5)
    rmass = d["rmass_inverse"].reshape(-1, 2)
    for prec in (4, 8):
        o = oracle.Oracle2D(mesh, prec, 1, cfg["dt"])
        ibool, midx, c2m = o.numbering()
        mi = o.mass_inverse()
        ours = mi[:, midx[ibool]]                 # (2, nspec, 5, 5) compute order
        ref = rmass[ref_ibool[c2m]].transpose(3, 0, 1, 2)
        err = np.abs(ours - ref).sum() / np.abs(ref).sum()
        assert err < cfg["tolerance"], (prec, err)


TRACE_CASES = ["elastic_free_80x60", "elastic_stacey_80x60", "acoustic_free_80x60", "acoustic_dirichlet_80x60",
               "acoustic_composite_80x60", "acoustic_stacey_80x60", "coupled_acoustic_top",
               "coupled_acoustic_bottom"]


@pytest.mark.parametrize("case", TRACE_CASES)
def test_forward_traces_match_specfem2d(case):
    """SPECFEM2D traces of the reference's Newmark tests
    (tests/unit-tests/displacement_tests/Newmark/newmark_tests.cpp:300-346,
    relative L1 per station <= 1e-3): elastic, acoustic with a free surface
    on all sides (Dirac source; displacement, velocity, acceleration and
    pressure traces), acoustic with Stacey, with Stacey and a free surface
    (composite elements), and the two coupled elastic-acoustic meshes.  The
    coupled cases run the FP64 restatement (the FP32 one agrees to 2e-5 and
    takes four times as long)."""
    mesh, cfg = load_case(case)
    traces = np.load(os.path.join(GOLDEN, case, "traces.npz"))
    prec = 8 if case.startswith("coupled") else 4
    o = oracle.Oracle2D(mesh, prec, 1, cfg["dt"])
    types = [TYPES[k] for k in cfg["trace_kinds"]]
    seis, t0 = o.forward(cfg["sources"], cfg["stations"], cfg["nstep"], cfg["dt"], 1, types)
    time = t0 + np.arange(cfg["nstep"]) * cfg["dt"]
    assert np.abs(time - traces["time"]).max() < 1e-3
    for it, k in enumerate(cfg["trace_kinds"]):
        ours = seis[:, it].transpose(1, 2, 0)
        if k == "p":
            ours = ours[:, :1]
        err = rel_l1_per_station(ours, traces[k])
        assert err.max() < cfg["tolerance"], (k, err.max())
        # a dropped or sign-flipped term (a coupling normal of the wrong sign
        # gives 0.45 on the coupled mesh) would be far outside this
        if k == "d":
            assert err.max() < 2e-4, (k, err.max())


def _geometry_case():
    import json

    from conftest import GOLDEN, load_case

    mesh, cfg = load_case("geometry_80x60")
    ref = np.load(os.path.join(GOLDEN, "geometry_80x60", "partial_derivatives.npz"))
    return mesh, cfg, ref


@pytest.mark.parametrize("precision", [8, 4])
def test_partial_derivatives_match_specfem2d(precision):
    """xix, xiz, gammax, gammaz and the jacobian of every GLL point of the
    reference's 80 x 60 test database against SPECFEM2D's arrays
    (tests/unit-tests/compute/partial_derivatives/compute_partial_derivatives_tests.cpp:
    absolute 1e-2 per value).  The restatement is also held to a relative bar:
    the largest error relative to the largest magnitude of its pair (xix with
    xiz, gammax with gammaz; xiz and gammax vanish on rectangles).  FP64 agrees
    to rounding (~1e-13).  FP32 sums shape-function derivatives times float
    control-node coordinates as the reference does in type_real
    (src/jacobian/jacobian.cpp, the ThreadVectorRange reductions): 4 km
    coordinates against 50 m elements cancel to ~3e-5 relative, 3e-2 absolute
    on jacobians near 940, so the fixture's absolute 1e-2 holds for FP64 and
    the relative bar 1e-4 for FP32."""
    mesh, cfg, ref = _geometry_case()
    o = oracle.Oracle2D(mesh, precision, 1, 1e-3)
    _, arrs = o.geometry()
    _, _, c2m = o.numbering()
    rel_bar = 1e-12 if precision == 8 else 1e-4
    scale = {"xix": ("xix", "xiz"), "xiz": ("xix", "xiz"), "gammax": ("gammax", "gammaz"),
             "gammaz": ("gammax", "gammaz"), "jacobian": ("jacobian",)}
    for name, ours in zip(("xix", "xiz", "gammax", "gammaz", "jacobian"), arrs):
        expect = ref[name][c2m]  # fixture is in mesh order, the oracle's in compute order
        err = np.abs(ours - expect).max()
        rel = err / max(np.abs(ref[k]).max() for k in scale[name])
        print(f"[geometry] fp{8 * precision} {name}: max abs {err:.3e}, rel {rel:.3e}")
        if precision == 8:
            assert err < cfg["partial_derivatives_tolerance"]
        assert rel < rel_bar


def test_locate_point_known_answers():
    """The five points of tests/unit-tests/algorithms/locate.cpp:37-58 on the
    same database: the element (mesh numbering) exactly, xi and gamma within
    1e-4, from the oracle and from the library's host assembly."""
    from specfempp_amd import Engine2D

    mesh, cfg, _ = _geometry_case()
    kat = cfg["locate_known_answers"]
    o = oracle.Oracle2D(mesh, 4, 1, 1e-3)
    _, _, c2m = o.numbering()
    pts = [dict(x=x, z=z) for x, z in kat["points"]]
    eng = Engine2D(mesh, [dict(x=2000.0, z=1500.0, f0=10.0, factor=1.0)], pts, 1e-3, 2, host_only=True)
    ispec, loc = eng.locations("receivers")
    eng.close()
    for k, p in enumerate(pts):
        e, xi, ga = o.locate(p["x"], p["z"])
        print(f"[locate] {p}: oracle mesh element {c2m[e]} xi {xi:.5f} gamma {ga:.5f}; "
              f"known {kat['mesh_ispec'][k]} {kat['xi'][k]} {kat['gamma'][k]}")
        assert c2m[e] == kat["mesh_ispec"][k]
        assert abs(xi - kat["xi"][k]) <= kat["tolerance_local"]
        assert abs(ga - kat["gamma"][k]) <= kat["tolerance_local"]
        assert ispec[k] == e
        assert abs(loc[k, 0] - xi) < 1e-7 and abs(loc[k, 1] - ga) < 1e-7
