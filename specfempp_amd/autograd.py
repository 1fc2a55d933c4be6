"""Differentiable 2-D forward model: per-point (rho, kappa, mu) -> displacement
seismograms, with the gradient from the engine's adjoint run.

``SeismogramModel`` wraps an :class:`Engine2D` whose run samples displacement
at every step (``nstep_between_samples = 1``, "displacement" first among the
seismogram types).  Its ``torch.autograd.Function`` runs ``forward_record`` in
the forward pass; the backward pass turns the incoming gradient dL/du(t) into
the adjoint source (dL/du(t) / dt, the form the kernels' dt factor expects,
frechet_derivative.hpp:110-118), runs the combined adjoint + backward loop and
converts the kernel densities into parameter gradients:

    dL/drho   = w J K_rho / rho
    dL/dkappa = -w J D,              D = -K_kappa / (lambda + mu)
    dL/dmu    = w J (2/3 D - 2 E),   E = -K_mu / (2 mu) + D / 3

where D = sum_t dt div(s+) div(s) and E = sum_t dt eps(s+) : eps(s) undo the
reference's kappa = lambda + mu and 1/3-deviatoric conventions
(frechet_derivative.hpp:33,78-92), and kappa is the bulk modulus the engine's
stress law uses (lambda = kappa - 2/3 mu).  Tensors live on the host; the
engine holds its own device memory.
"""
from __future__ import annotations

import numpy as np
import torch


class SeismogramModel:
    def __init__(self, engine, receivers, snapshot_every=0):
        if engine.dim != 2:
            raise ValueError("the differentiable model is implemented for 2-D engines")
        if engine.nsig != engine.nstep:
            raise ValueError("the engine must sample seismograms at every step")
        if engine.types[0] != "displacement":
            raise ValueError("displacement must be the first seismogram type")
        self.engine = engine
        self.receivers = list(receivers)
        self.dt = float(engine.dt)
        self.wj = engine.point_weights().astype(np.float64)
        self.snapshot_every = int(snapshot_every)

    def __call__(self, rho, kappa, mu):
        return _Seismograms.apply(rho, kappa, mu, self)

    def gradients(self, rho, kappa, mu, dl_du):
        """Parameter gradients for dL/du (nstep, nrec, 2); numpy in and out."""
        eng = self.engine
        stf = np.transpose(np.asarray(dl_du, dtype=np.float64), (1, 0, 2)) / self.dt
        eng.set_adjoint_sources(self.receivers, stf)
        eng.adjoint_run()
        K = eng.kernels()
        rho = np.asarray(rho, dtype=np.float64)
        kappa = np.asarray(kappa, dtype=np.float64)
        mu = np.asarray(mu, dtype=np.float64)
        lam = kappa - 2.0 / 3.0 * mu
        D = -K["kappa"].astype(np.float64) / (lam + mu)
        E = -K["mu"].astype(np.float64) / (2.0 * mu) + D / 3.0
        g_rho = self.wj * K["rho"].astype(np.float64) / rho
        g_kappa = -self.wj * D
        g_mu = self.wj * (2.0 / 3.0 * D - 2.0 * E)
        return g_rho, g_kappa, g_mu

    def forward_numpy(self, rho, kappa, mu):
        eng = self.engine
        eng.set_properties(rho, kappa, mu)
        eng.initialize()
        eng.forward_record(self.snapshot_every)
        return eng.seismograms()[:, 0].astype(np.float64)  # (nstep, nrec, 2)


class _Seismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, rho, kappa, mu, model):
        seis = model.forward_numpy(rho.detach().cpu().numpy(), kappa.detach().cpu().numpy(),
                                   mu.detach().cpu().numpy())
        ctx.model = model
        ctx.save_for_backward(rho, kappa, mu)
        return torch.from_numpy(seis).to(dtype=rho.dtype)

    @staticmethod
    def backward(ctx, grad):
        rho, kappa, mu = ctx.saved_tensors
        g = ctx.model.gradients(rho.detach().cpu().numpy(), kappa.detach().cpu().numpy(), mu.detach().cpu().numpy(),
                                grad.detach().cpu().numpy())
        return tuple(torch.from_numpy(x).to(dtype=rho.dtype) for x in g) + (None,)
