"""GPU parity of the 2-D combined adjoint + backward run against the oracle.

The oracle's combined loop is pinned on the CPU (test_oracle_adjoint.py:
time reversibility and finite differences of the misfit).  Here the HIP
engine's forward record, backward seismograms and the six kernels are
compared with the FP32 and FP64 oracle on the same inputs.  Bars: relative L2
<= 1e-5 against the FP32 oracle, or, where FP32 rounding alone moves the
result further (the kernels are sums over every step of products of two
wavefields), a distance to the FP64 oracle within 1.5x the FP32 oracle's own.
"""
import numpy as np
import pytest

import oracle
from conftest import rel_l2
from specfempp_amd import meshes

pytestmark = pytest.mark.gpu

N, H, DT, NSTEP = 16, 50.0, 1e-3, 500
TYPES = ("displacement",)
SRC = [dict(x=310.0, z=420.0, angle=30.0, f0=15.0, factor=1e10)]
REC = [dict(x=600.0, z=650.0), dict(x=180.0, z=200.0), dict(x=520.0, z=130.0)]
NAMES = ("rho", "mu", "kappa", "rhop", "alpha", "beta")


def _mesh():
    L = N * H
    return meshes.quad_mesh(N, N, 0, L, 0, L, materials=[meshes.elastic(2400.0, 3000.0, 1732.0)],
                            absorbing=("bottom", "right", "left", "top"))


def _report(name, **vals):
    print(f"[adjoint] {name}: " + ", ".join(f"{k}={v:.3e}" for k, v in vals.items()), flush=True)


def _check(e32, e64, floor):
    assert e32 <= 1e-5 or e64 <= 1.5 * floor, (e32, e64, floor)


@pytest.fixture(scope="module")
def oracle_runs():
    mesh = _mesh()
    out = {}
    rng = np.random.default_rng(7)
    # a seeded adjoint source (smooth random traces at the receivers)
    t = np.arange(NSTEP) * DT
    stf = np.zeros((len(REC), NSTEP, 2))
    for r in range(len(REC)):
        for c in range(2):
            a, f, ph = rng.uniform(0.5, 1.5), rng.uniform(8, 20), rng.uniform(0, 2 * np.pi)
            stf[r, :, c] = 1e-6 * a * np.sin(2 * np.pi * f * t + ph) * np.exp(-((t - 0.25) / 0.1) ** 2)
    for prec in (4, 8):
        for k in (0, 23):
            o = oracle.Oracle2D(mesh, prec, 1, DT)
            seis, _, fields, record, *snaps = o.forward_record(SRC, REC, NSTEP, DT, 1, TYPES, snapshot_every=k)
            kern, bseis = o.combined(SRC, REC, stf, REC, NSTEP, DT, fields, record, 1, TYPES, snapshot_every=k,
                                     snapshots=snaps[0] if snaps else None)
            out[prec, k] = (seis, kern, bseis)
    return mesh, stf, out


@pytest.mark.parametrize("scatter,snapshot_every", [("colored", 0), ("atomic", 0), ("colored", 23)])
def test_adjoint_kernels_match_oracle(oracle_runs, scatter, snapshot_every):
    from specfempp_amd import Engine2D

    mesh, stf, out = oracle_runs
    eng = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, TYPES, scatter=scatter)
    eng.initialize()
    eng.forward_record(snapshot_every)
    fwd = eng.seismograms().astype(np.float64)
    eng.set_adjoint_sources(REC, stf)
    eng.adjoint_run()
    bwd = eng.seismograms().astype(np.float64)
    kern = eng.kernels()
    eng.close()
    (s32, k32, b32), (s64, k64, b64) = out[4, snapshot_every], out[8, snapshot_every]
    scatter = f"{scatter}_k{snapshot_every}"
    e32 = rel_l2(fwd, s32)
    _report(f"{scatter}/forward", vs_oracle_fp32=e32, vs_oracle_fp64=rel_l2(fwd, s64))
    assert e32 <= 1e-5
    e32, e64, floor = rel_l2(bwd, b32), rel_l2(bwd, b64), rel_l2(b32, b64)
    _report(f"{scatter}/backward_seismograms", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
    _check(e32, e64, floor)
    for k, name in enumerate(NAMES):
        ours = kern[name].astype(np.float64)
        e32, e64, floor = rel_l2(ours, k32[k]), rel_l2(ours, k64[k]), rel_l2(k32[k], k64[k])
        _report(f"{scatter}/K_{name}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)


def test_adjoint_requires_recorded_forward():
    from specfempp_amd import Engine2D
    from specfempp_amd.engine import SfxError

    eng = Engine2D(_mesh(), SRC, REC, DT, 50, 1, TYPES)
    eng.initialize()
    eng.set_adjoint_sources(REC, np.zeros((len(REC), 50, 2)))
    with pytest.raises(SfxError, match="sfx_forward_record"):
        eng.adjoint_run()
    eng.forward_record()
    eng.adjoint_run()
    with pytest.raises(SfxError, match="sfx_forward_record"):
        eng.adjoint_run()  # the recorded state is consumed
    eng.close()


def test_autograd_gradient_matches_oracle_and_finite_differences():
    """torch.autograd through the engine: the gradient of the L2 waveform
    misfit with respect to per-point (rho, kappa, mu) equals the gradient from
    the FP64 oracle's kernels, and predicts central finite differences of the
    FP64 oracle's misfit along a smooth direction."""
    import torch

    from specfempp_amd import Engine2D
    from specfempp_amd.autograd import SeismogramModel

    mesh = _mesh()
    o = oracle.Oracle2D(mesh, 8, 1, DT)
    rho0, kap0, mu0, wj = o.properties()
    ex, ez = np.arange(mesh.nspec) % N, np.arange(mesh.nspec) // N
    gl = np.array([-1.0, -np.sqrt(3 / 7), 0.0, np.sqrt(3 / 7), 1.0])
    X = (ex[:, None, None] + (gl[None, None, :] + 1) / 2 + 0 * gl[None, :, None]) * H
    Z = (ez[:, None, None] + (gl[None, :, None] + 1) / 2 + 0 * gl[None, None, :]) * H
    interior = ((ex >= 2) & (ex < N - 2) & (ez >= 2) & (ez < N - 2))[:, None, None]
    bump = np.exp(-((X - 420) ** 2 + (Z - 380) ** 2) / 120.0 ** 2) * interior
    o_true = oracle.Oracle2D(mesh, 8, 1, DT)
    o_true.set_properties(rho0 * (1 + 0.02 * bump), kap0 * (1 + 0.02 * bump), mu0 * (1 - 0.02 * bump))
    data = torch.from_numpy(o_true.forward(SRC, REC, NSTEP, DT, 1, TYPES)[0][:, 0])

    eng = Engine2D(mesh, SRC, REC, DT, NSTEP, 1, TYPES)
    model = SeismogramModel(eng, REC)
    params = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in (rho0, kap0, mu0)]
    u = model(*params)
    loss = 0.5 * torch.sum((u - data) ** 2) * DT
    loss.backward()
    grads = [p.grad.numpy() for p in params]
    eng.close()

    # the same gradient from the FP64 oracle's kernels
    seis, _, fields, record = o.forward_record(SRC, REC, NSTEP, DT, 1, TYPES)
    res = seis[:, 0] - data.numpy()
    K = o.combined(SRC, REC, np.transpose(res, (1, 0, 2)), REC, NSTEP, DT, fields, record, 1, TYPES)[0]
    lam = kap0 - 2.0 / 3.0 * mu0
    D = -K[2] / (lam + mu0)
    E = -K[1] / (2 * mu0) + D / 3.0
    ref = (wj * K[0] / rho0, -wj * D, wj * (2.0 / 3.0 * D - 2.0 * E))
    for name, g, r in zip(("rho", "kappa", "mu"), grads, ref):
        err = rel_l2(g, r)
        _report(f"autograd/d_{name}", vs_oracle_fp64=err)
        assert err <= 1e-3

    # directional derivative vs central differences of the FP64 misfit
    g = np.exp(-((X - 380) ** 2 + (Z - 450) ** 2) / 100.0 ** 2) * interior
    eps = 1e-3

    def misfit(rho, kap, mu):
        oo = oracle.Oracle2D(mesh, 8, 1, DT)
        oo.set_properties(rho, kap, mu)
        s = oo.forward(SRC, REC, NSTEP, DT, 1, TYPES)[0][:, 0]
        return 0.5 * np.sum((s - data.numpy()) ** 2) * DT

    dirs = {"rho": (rho0 * g, 0 * g, 0 * g), "kappa": (0 * g, kap0 * g, 0 * g), "mu": (0 * g, 0 * g, mu0 * g)}
    for name, (dr, dk, dm) in dirs.items():
        fd = (misfit(rho0 + eps * dr, kap0 + eps * dk, mu0 + eps * dm) -
              misfit(rho0 - eps * dr, kap0 - eps * dk, mu0 - eps * dm)) / 2
        pred = eps * sum(np.sum(gr * d) for gr, d in zip(grads, (dr, dk, dm)))
        rel = abs(pred - fd) / abs(fd)
        _report(f"autograd/fd_{name}", fd=fd, autograd=pred, rel=rel)
        assert rel <= 1e-3


# ---- configuration 5, reduced: the adjoint workload at 100^2 -----------------
# Same 20 m elements, 400 m checkerboard (+-5 %), 8 seeded sources, 100
# receivers, dt 5e-4 and checkpoints every 55 steps as the full 200^2 x 3000
# configuration; 100^2 elements and 1500 steps keep the FP32/FP64 oracle runs
# (in parallel worker processes) within about a minute.
C5_N, C5_NSTEP, C5_SNAP, C5_EPS = 100, 1500, 55, 1e-4


@pytest.fixture(scope="module")
def config5_reduced():
    import _oracle_jobs as jobs

    true, start = meshes.config5(n=C5_N, nstep=C5_NSTEP)
    args = (start.sources, start.receivers, C5_NSTEP, start.dt)
    with jobs.pool(2) as ex:
        f_data = ex.submit(jobs.forward_2d, true.mesh, 8, *args)
        f_start = ex.submit(jobs.forward_2d, start.mesh, 8, *args)
        data = f_data.result()[:, 0]
        s64 = f_start.result()[:, 0]
    # one adjoint source for engine and oracle alike: the FP64 residual at the
    # receivers, (nrec, nstep, 2) in forward time order
    stf = np.ascontiguousarray(np.transpose(s64 - data, (1, 0, 2)))
    return true, start, data, stf


def _stacey_ring_mask(n):
    """Points of elements off the absorbing ring (all four sides absorb).  The
    reference's Fréchet kernels carry no derivative of the Stacey impedances
    or of the mass matrix's absorbing term (frechet_derivative.hpp:14-138), so
    the finite-difference directions leave the boundary elements alone."""
    ez, ex = np.divmod(np.arange(n * n), n)
    inner = (ex > 0) & (ex < n - 1) & (ez > 0) & (ez < n - 1)
    return np.broadcast_to(inner[:, None, None], (n * n, 5, 5))


def test_config5_reduced_kernels_match_oracle_and_finite_differences(config5_reduced):
    """Forward record, backward seismograms and all six kernels against the
    FP32/FP64 oracle (time_marching.tpp:126-247, frechet_derivative.hpp:14-138);
    then the torch.autograd gradient of the FP64 L2 misfit against central
    finite differences of the FP64 oracle's misfit along 4 of the 16 seeded
    random perturbations (N(0, 1) per GLL point on rho, mu, kappa; seeds 5.1 to
    5.4; eps = 1e-4 relative), bar 1e-3."""
    import torch

    import _oracle_jobs as jobs
    from specfempp_amd import Engine2D
    from specfempp_amd.autograd import SeismogramModel

    true, start, data, stf = config5_reduced
    dt, n = start.dt, C5_N
    print(f"[case] config 5 reduced {n}^2 x {C5_NSTEP} steps, snapshots every {C5_SNAP}", flush=True)
    o64 = oracle.Oracle2D(start.mesh, 8, 1, dt)
    rho0, kap0, mu0, _ = o64.properties()
    mask = _stacey_ring_mask(n)
    dirs = []
    for i in range(1, 5):
        rng = np.random.default_rng([5, i])
        dirs.append([rng.standard_normal(rho0.shape) * mask for _ in range(3)])  # rho, mu, kappa
    args = (start.sources, start.receivers, C5_NSTEP, dt)
    with jobs.pool(min(10, jobs.host_workers())) as ex:
        f_adj = {prec: ex.submit(jobs.adjoint_2d, start.mesh, prec, *args, stf, C5_SNAP) for prec in (4, 8)}
        f_fd = []
        for d_rho, d_mu, d_kap in dirs:
            for sgn in (1.0, -1.0):
                props = (rho0 * (1 + sgn * C5_EPS * d_rho), kap0 * (1 + sgn * C5_EPS * d_kap),
                         mu0 * (1 + sgn * C5_EPS * d_mu))
                f_fd.append(ex.submit(jobs.misfit_2d, start.mesh, *args, data, props))

        # engine: kernels for the same adjoint source
        eng = Engine2D(start.mesh, start.sources, start.receivers, dt, C5_NSTEP, 1, TYPES, scatter="colored")
        eng.initialize()
        eng.forward_record(C5_SNAP)
        fwd = eng.seismograms().astype(np.float64)
        eng.set_adjoint_sources(start.receivers, stf)
        eng.adjoint_run()
        bwd = eng.seismograms().astype(np.float64)
        kern = eng.kernels()
        # engine: gradient of the misfit through torch.autograd
        model = SeismogramModel(eng, start.receivers, snapshot_every=C5_SNAP)
        params = [torch.tensor(x, dtype=torch.float64, requires_grad=True) for x in (rho0, kap0, mu0)]
        u = model(*params)
        loss = 0.5 * torch.sum((u - torch.from_numpy(data)) ** 2) * dt
        loss.backward()
        g_rho, g_kap, g_mu = [p.grad.numpy() for p in params]
        eng.close()

        (s32, k32, b32), (s64, k64, b64) = f_adj[4].result(), f_adj[8].result()
        e32 = rel_l2(fwd, s32)
        _report("config5_reduced/forward", vs_oracle_fp32=e32, vs_oracle_fp64=rel_l2(fwd, s64))
        assert e32 <= 1e-5
        e32, e64, floor = rel_l2(bwd, b32), rel_l2(bwd, b64), rel_l2(b32, b64)
        _report("config5_reduced/backward_seismograms", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
        _check(e32, e64, floor)
        for k, name in enumerate(NAMES):
            ours = kern[name].astype(np.float64)
            e32, e64, floor = rel_l2(ours, k32[k]), rel_l2(ours, k64[k]), rel_l2(k32[k], k64[k])
            _report(f"config5_reduced/K_{name}", vs_oracle_fp32=e32, vs_oracle_fp64=e64, fp32_floor=floor)
            _check(e32, e64, floor)
            assert e64 <= 1e-4  # the north-star bar against the reference kernels (FP64 oracle)

        J = [f.result() for f in f_fd]
    for i, (d_rho, d_mu, d_kap) in enumerate(dirs):
        fd = (J[2 * i] - J[2 * i + 1]) / 2.0
        pred = C5_EPS * float(np.sum(g_rho * rho0 * d_rho + g_kap * kap0 * d_kap + g_mu * mu0 * d_mu))
        rel = abs(pred - fd) / abs(fd)
        _report(f"config5_reduced/fd_seed5.{i + 1}", fd=fd, autograd=pred, rel=rel)
        assert rel <= 1e-3
