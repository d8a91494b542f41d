"""Differentiable frame-averaged solute-solvent energy.

``frame_averaged_energy(engine, frames, q, eps, sigma, configs)`` returns
<u> = sum_f sum_b w u / sum_f sum_b w, the weighted average over frames and
solute configurations of u = sum_{k>=1} uvengy(k) (the solute-solvent
energy get_uv_energy, engproc.F90:662-756, distributes over molecules), as a
differentiable function of the solute's per-site charges q (e), LJ epsilon
(kcal/mol) and sigma (Angstrom).  The forward pass and the gradient come
from one call per frame of the hand-written adjoint kernels
(``erm_solute_gradient``); the backward pass scales the stored gradient.
The energy-distribution histograms are not touched.
"""
from __future__ import annotations

import numpy as np
import torch

from .system import CHARGE_SCALE


def _host(t) -> np.ndarray:
    return t.detach().to("cpu", torch.float64).numpy().copy()


class FrameAveragedEnergy(torch.autograd.Function):
    """autograd.Function over (q, eps, sigma); see ``frame_averaged_energy``."""

    @staticmethod
    def forward(ctx, q, eps, sigma, engine, frames, configs, cmbrule):
        qs = _host(q) * CHARGE_SCALE
        ep = _host(eps)
        sg = _host(sigma)
        sw = swu = 0.0
        gq = np.zeros_like(qs)
        ge = np.zeros_like(ep)
        gs = np.zeros_like(sg)
        for frame, conf in zip(frames, configs):
            pos, cell, weight = frame
            if isinstance(pos, torch.Tensor) and pos.is_cuda:
                if pos.dtype != torch.float64 or not pos.is_contiguous():
                    raise ValueError("device coordinates must be contiguous float64")
                # the library copies the coordinates on its own stream: wait
                # for whatever torch stream produced them first
                torch.cuda.current_stream(pos.device).synchronize()
                engine.set_frame_device(pos.data_ptr(), cell, weight)
            else:
                engine.set_frame(np.asarray(pos, dtype=np.float64), cell, weight)
            a = engine.solute_gradient(cmbrule=cmbrule, q=qs, eps=ep, sigma=sg, **conf)
            sw += a[0]
            swu += a[1]
            gq += a[2]
            ge += a[3]
            gs += a[4]
        if sw == 0.0:
            raise ValueError("no configurations (total weight 0)")
        dev = q.device
        ctx.save_for_backward(
            torch.as_tensor(gq / sw * CHARGE_SCALE, device=dev).to(q.dtype),
            torch.as_tensor(ge / sw, device=dev).to(eps.dtype),
            torch.as_tensor(gs / sw, device=dev).to(sigma.dtype))
        return torch.as_tensor(swu / sw, dtype=torch.float64, device=dev)

    @staticmethod
    def backward(ctx, g):
        gq, ge, gs = ctx.saved_tensors
        return g * gq, g * ge, g * gs, None, None, None, None


def frame_averaged_energy(engine, frames, q, eps, sigma, configs, cmbrule: int = 0):
    """<u> as a differentiable function of the solute's per-site parameters.

    engine   an ``ermod_amd.api.ErmGPU`` whose topology holds the solute
    frames   iterable of (sitepos, cell, weight): sitepos a host array or a
             contiguous float64 device tensor of natom x 3 coordinates
    q, eps, sigma   1-D tensors over the solute's sites (e, kcal/mol, A)
    configs  one dict per frame of ``ErmGPU.solute_gradient`` arguments:
             {"solutes": [...]} (soln), {"solute_pos": ...} or
             {"n": N, "seed": s, "first_index": i} (test particle)
    cmbrule  0 arithmetic / 1 geometric sigma combination (setconf.F90:810-817)
    """
    frames = list(frames)
    configs = list(configs)
    if len(frames) != len(configs):
        raise ValueError("one configuration set per frame")
    return FrameAveragedEnergy.apply(q, eps, sigma, engine, frames, configs, cmbrule)
