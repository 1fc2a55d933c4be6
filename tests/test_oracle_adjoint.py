"""Oracle pinning of the combined adjoint + backward loop (CPU).

The reference's Fréchet kernels have no golden artifact, so the restatement
of time_marching.tpp:126-247 and frechet_derivative.hpp:14-138 is pinned by
what the mathematics fixes:
- the backward field, stepped with -dt from the forward run's final state and
  replaying the recorded Stacey-element forces, retraces the forward
  seismograms (Newmark time reversibility);
- the kernels predict the change of the L2 waveform misfit under smooth model
  perturbations, checked against central finite differences of forward runs
  (rho; mu at fixed lambda and lambda at fixed mu, both derived from the
  reference's kappa and mu kernels with its kappa = lambda + mu and
  1/3-deviatoric conventions);
- with a zero adjoint source every kernel is zero.
"""
import numpy as np
import pytest

import oracle
from specfempp_amd import meshes

N, H, DT, NSTEP = 16, 50.0, 1e-3, 500
TYPES = ("displacement",)
SRC = [dict(x=310.0, z=420.0, angle=30.0, f0=15.0, factor=1e10)]
REC = [dict(x=600.0, z=650.0), dict(x=180.0, z=200.0), dict(x=520.0, z=130.0)]


def _mesh():
    L = N * H
    return meshes.quad_mesh(N, N, 0, L, 0, L, materials=[meshes.elastic(2400.0, 3000.0, 1732.0)],
                            absorbing=("bottom", "right", "left", "top"))


def _point_coords(nspec):
    """GLL point coordinates (nspec, 5, 5) of the structured mesh, mesh order."""
    ex, ez = np.arange(nspec) % N, np.arange(nspec) // N
    gl = np.array([-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0])
    X = (ex[:, None, None] + (gl[None, None, :] + 1) / 2 + 0 * gl[None, :, None]) * H
    Z = (ez[:, None, None] + (gl[None, :, None] + 1) / 2 + 0 * gl[None, None, :]) * H
    interior = ((ex >= 2) & (ex < N - 2) & (ez >= 2) & (ez < N - 2))[:, None, None]
    return X, Z, interior


@pytest.fixture(scope="module")
def forward64():
    mesh = _mesh()
    o = oracle.Oracle2D(mesh, 8, 1, DT)
    seis, _, fields, record = o.forward_record(SRC, REC, NSTEP, DT, 1, TYPES)
    return mesh, o, seis[:, 0], fields, record


@pytest.mark.parametrize("precision,bar", [(8, 1e-12), (4, 1e-4)])
def test_backward_field_retraces_forward(precision, bar):
    mesh = _mesh()
    o = oracle.Oracle2D(mesh, precision, 1, DT)
    seis, _, fields, record = o.forward_record(SRC, REC, NSTEP, DT, 1, TYPES)
    kern, bseis = o.combined(SRC, REC, np.zeros((len(REC), NSTEP, 2)), REC, NSTEP, DT, fields, record, 1, TYPES)
    fw, bw = seis[:, 0], bseis[::-1, 0]
    err = np.linalg.norm(fw - bw) / np.linalg.norm(fw)
    print(f"[adjoint] reversibility precision={precision}: {err:.3e}")
    assert err <= bar
    assert np.abs(kern).max() == 0.0


def test_checkpoints_bound_fp32_drift():
    """Resetting the backward field to forward checkpoints every k steps keeps
    the FP32 reconstruction closer to the forward run than pure time reversal."""
    mesh = _mesh()
    o = oracle.Oracle2D(mesh, 4, 1, DT)
    seis, _, fields, record, snaps = o.forward_record(SRC, REC, NSTEP, DT, 1, TYPES, snapshot_every=23)
    assert snaps.shape == ((NSTEP - 2) // 23 + 1, 3, 2, o.nglob)
    zero = np.zeros((len(REC), NSTEP, 2))
    _, b_pure = o.combined(SRC, REC, zero, REC, NSTEP, DT, fields, record, 1, TYPES)
    _, b_snap = o.combined(SRC, REC, zero, REC, NSTEP, DT, fields, record, 1, TYPES, snapshot_every=23,
                           snapshots=snaps)
    fw = seis[:, 0]
    e_pure = np.linalg.norm(fw - b_pure[::-1, 0]) / np.linalg.norm(fw)
    e_snap = np.linalg.norm(fw - b_snap[::-1, 0]) / np.linalg.norm(fw)
    print(f"[adjoint] FP32 reconstruction: pure {e_pure:.3e}, checkpoints every 23 steps {e_snap:.3e}")
    assert e_snap < 0.5 * e_pure


def _misfit(mesh, rho, kappa, mu, data):
    o = oracle.Oracle2D(mesh, 8, 1, DT)
    o.set_properties(rho, kappa, mu)
    s, _ = o.forward(SRC, REC, NSTEP, DT, 1, TYPES)
    return 0.5 * np.sum((s[:, 0] - data) ** 2) * DT


def test_kernels_match_finite_differences(forward64):
    mesh, o, u, fields, record = forward64
    rho0, kap0, mu0, wj = o.properties()
    X, Z, interior = _point_coords(mesh.nspec)
    # data from a model with a smooth anomaly away from the absorbing edges
    bump = np.exp(-((X - 420) ** 2 + (Z - 380) ** 2) / 120.0 ** 2) * interior
    o_true = oracle.Oracle2D(mesh, 8, 1, DT)
    o_true.set_properties(rho0 * (1 + 0.02 * bump), kap0 * (1 + 0.02 * bump), mu0 * (1 - 0.02 * bump))
    data = o_true.forward(SRC, REC, NSTEP, DT, 1, TYPES)[0][:, 0]
    res = u - data
    kern, _ = o.combined(SRC, REC, np.transpose(res, (1, 0, 2)), REC, NSTEP, DT, fields, record, 1, TYPES)
    Krho, Kmu, Kkap, Krhop, Kalpha, Kbeta = kern
    lam = kap0 - 2.0 / 3.0 * mu0
    kap2 = lam + mu0                       # the reference's kappa (frechet_derivative.hpp:33)
    D = -Kkap / kap2                       # sum_t dt div(s+) div(s)
    E = -Kmu / (2 * mu0) + D / 3.0         # sum_t dt eps(s+) : eps(s)
    g = np.exp(-((X - 380) ** 2 + (Z - 450) ** 2) / 100.0 ** 2) * interior
    eps = 1e-3
    cases = {
        "rho": (lambda s: (rho0 * (1 + s * eps * g), kap0, mu0), np.sum(wj * Krho * g) * eps),
        "mu_fixed_lambda": (lambda s: (rho0, lam + 2.0 / 3.0 * mu0 * (1 + s * eps * g), mu0 * (1 + s * eps * g)),
                            np.sum(wj * (-2 * E) * mu0 * g) * eps),
        "lambda": (lambda s: (rho0, lam * (1 + s * eps * g) + 2.0 / 3.0 * mu0, mu0),
                   np.sum(wj * (-D) * lam * g) * eps),
    }
    for name, (model, pred) in cases.items():
        fd = (_misfit(mesh, *model(1), data) - _misfit(mesh, *model(-1), data)) / 2
        rel = abs(pred - fd) / abs(fd)
        print(f"[adjoint] {name}: fd={fd:.6e} kernel={pred:.6e} rel={rel:.2e}")
        assert rel < 1e-3
    # the derived kernels are the reference's combinations (frechet_derivative.hpp:120-135)
    # (summed per step, so equal up to the rounding of the step sums)
    def close(a, b):
        return np.abs(a - b).max() <= 1e-9 * np.abs(b).max()
    assert close(Krhop, Krho + Kkap + Kmu)
    assert close(Kbeta, 2 * (Kmu - 4.0 / 3.0 * mu0 / kap2 * Kkap))
    assert close(Kalpha, 2 * (1 + 4.0 / 3.0 * mu0 / kap2) * Kkap)
