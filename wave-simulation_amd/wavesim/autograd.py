"""torch.autograd binding of the adjoint-state gradient (2D elastic).

``elastic_seismograms(vp, vs, rho, solver)`` returns the seismograms of every
shot of a prepared solver handle, concatenated over shots ([sum nrec, NT],
float32); its backward pass runs the time-reversed adjoint on the GPU
(checkpointed recomputation, FP64 gradient accumulation) and returns
dJ/dvp, dJ/dvs, dJ/drho for whatever scalar J the caller builds from the
seismograms.  torch only carries the arrays; all work is in libwavehip.so.
"""
from __future__ import annotations

import numpy as np
import torch


def _host(t):
    return np.ascontiguousarray(t.detach().to("cpu", torch.float32).numpy().reshape(-1))


class ElasticSeismograms(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vp, vs, rho, solver, every, fp16):
        solver.prepare_for_modelling(_host(vp), _host(vs), _host(rho))
        solver.set_checkpointing(every, fp16)
        solver.run(0, solver.cfg.NT)
        seis = np.concatenate([solver.seismograms(k) for k in range(solver.cfg.nshots)], axis=0)
        ctx.solver = solver
        ctx.shape = vp.shape
        ctx.device = vp.device
        return torch.from_numpy(seis).to(vp.device)

    @staticmethod
    def backward(ctx, grad_seis):
        sol = ctx.solver
        g = np.ascontiguousarray(grad_seis.detach().to("cpu", torch.float32).numpy())
        off = 0
        for k in range(sol.cfg.nshots):
            n = sol._nrec.get(k, 0)
            sol.set_adjoint_source(k, g[off:off + n])
            off += n
        sol.run_adjoint()
        out = [torch.from_numpy(sol.gradient(p).astype(np.float32)).reshape(ctx.shape).to(ctx.device)
               for p in ("vp", "vs", "rho")]
        return out[0], out[1], out[2], None, None, None


def elastic_seismograms(vp, vs, rho, solver, every=50, fp16=True):
    """Differentiable forward model: (vp, vs, rho) -> seismograms of all shots of `solver`."""
    return ElasticSeismograms.apply(vp, vs, rho, solver, every, fp16)
