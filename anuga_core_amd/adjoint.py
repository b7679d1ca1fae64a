"""Differentiable DE1 runs for calibration: a torch.autograd.Function whose
forward is a fixed-timestep run on the device and whose backward is the
library's discrete adjoint (anuga_adjoint_record / anuga_adjoint_gradient).

    run = GaugeRun(domain, initial, nsteps=2000, dt=0.05, gauges=g, obs_every=10)
    n = torch.full((N,), 0.03, dtype=torch.float64, requires_grad=True)
    z = torch.tensor(bed, requires_grad=True)
    stage = run(n, z)                       # (nsteps // obs_every, len(g))
    loss = ((stage - observed) ** 2).sum()
    loss.backward()                         # n.grad, z.grad per triangle

Each backward differentiates the run of its own forward: when other runs were
recorded on the domain in between (two terms of one loss, a closure
optimiser), it records its run again before the adjoint sweep.

The gradient is that of the discrete scheme the forward computes (FP64 path,
branches of the limiter and of wet/dry taken where the forward took them);
the bed derivative is at fixed stage. Tensors are float64 on the CPU: the
state stays on the device inside the library, only the per-triangle
parameters and the gauge series cross.
"""

from __future__ import annotations

import numpy as np


class GaugeRun:
    """A fixed-timestep DE1 run of `domain` from `initial` (centroid stage,
    xmomentum, ymomentum arrays) as a function of friction and bed."""

    def __init__(self, domain, initial, nsteps, dt, gauges, obs_every=1, checkpoint_every=100):
        import torch

        self.domain = domain
        self.initial = {k: np.ascontiguousarray(v, np.float64) for k, v in initial.items()}
        self.nsteps, self.dt = int(nsteps), float(dt)
        self.gauges = np.ascontiguousarray(gauges, np.int64)
        self.obs_every, self.checkpoint_every = int(obs_every), int(checkpoint_every)
        run = self

        class _Fn(torch.autograd.Function):
            @staticmethod
            def forward(ctx, friction, elevation):
                n = friction.detach().cpu().numpy().copy()
                z = elevation.detach().cpu().numpy().copy()
                out = torch.from_numpy(run.forward(n, z))
                # the run this output came from, and its parameters, so that a
                # backward after other forwards re-records this one
                ctx.record = run.domain.adjoint_record_serial
                ctx.params = (n, z)
                return out

            @staticmethod
            def backward(ctx, grad):
                if run.domain.adjoint_record_serial != ctx.record:
                    run.forward(*ctx.params)
                gn, gz, _ = run.domain.adjoint_gradient(grad.detach().cpu().numpy())
                return torch.from_numpy(gn), torch.from_numpy(gz)

        self._fn = _Fn

    def forward(self, friction, elevation):
        """The gauge stage series (numpy) for these parameters."""
        d = self.domain
        d.set_quantity("friction", np.asarray(friction, np.float64), location="centroids")
        d.set_quantity("elevation", np.asarray(elevation, np.float64), location="centroids")
        for name in ("stage", "xmomentum", "ymomentum"):
            d.set_quantity(name, self.initial[name], location="centroids")
        d.set_time(d.starttime)
        return d.adjoint_record(self.nsteps, self.dt, self.gauges, self.obs_every,
                                self.checkpoint_every)

    def __call__(self, friction, elevation):
        return self._fn.apply(friction, elevation)
