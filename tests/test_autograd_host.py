"""Host logic of the differentiable energy (CPU).

``frame_averaged_energy`` is driven here by an engine stand-in whose
solute_gradient is the oracle's, so unit conversion (charges in e), frame
weighting and the backward pass are checked without a GPU: the autograd
gradient must equal central differences of the forward value.
"""
import numpy as np
import pytest
import torch

import oracle
from ermod_amd import synth
from ermod_amd.autograd import frame_averaged_energy
from ermod_amd.system import CHARGE_SCALE, site_lj


class OracleEngine:
    def __init__(self, system):
        self.o = oracle.Oracle(system)
        self.system = system
        self.w = 1.0

    def set_frame(self, pos, cell, weight):
        self.o.set_frame(pos, cell, weight)
        self.w = weight

    def solute_gradient(self, cmbrule=0, q=None, eps=None, sigma=None, solutes=None, solute_pos=None):
        slt = int(self.system.solutes()[0])
        ns = self.system.nsite(slt)
        confs = [(m, None) for m in solutes] if solutes is not None else \
            [(slt, p) for p in np.asarray(solute_pos).reshape(-1, ns, 3)]
        acc = [0.0, 0.0, np.zeros(ns), np.zeros(ns), np.zeros(ns)]
        for m, p in confs:
            u, gq, ge, gs = self.o.solute_gradient(m, p, cmbrule, q=q, eps=eps, sigma=sigma)
            acc[0] += self.w
            acc[1] += self.w * u
            acc[2] += self.w * gq
            acc[3] += self.w * ge
            acc[4] += self.w * gs
        return acc


def test_autograd_matches_finite_differences():
    syn = synth.tiny_soln()
    s = syn.system
    slt = syn.solute_mol
    b, e = s.mol_begin[slt], s.mol_begin[slt + 1]
    ep, sg = site_lj(s)
    q = torch.tensor(s.charge[b:e] / CHARGE_SCALE, dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(ep[b:e], dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(sg[b:e], dtype=torch.float64, requires_grad=True)
    frames = [(syn.frame(i), syn.cell, w) for i, w in ((0, 1.0), (1, 0.5), (2, 2.0))]
    confs = [{"solutes": [slt]}] * 3
    eng = OracleEngine(s)
    u = frame_averaged_energy(eng, frames, q, eps, sig, confs)
    u.backward()
    # weighted average of the per-frame energies
    us = []
    for f, _, w in frames:
        eng.o.set_frame(f, syn.cell, w)
        us.append(eng.o.solute_gradient(slt, None, 0)[0])
    ws = np.array([1.0, 0.5, 2.0])
    assert float(u.detach()) == pytest.approx(float(np.dot(ws, us) / ws.sum()), rel=1e-12)
    h = 1e-4
    for name, t in (("q", q), ("eps", eps), ("sigma", sig)):
        for i in range(len(t)):
            vals = []
            for sgn in (1, -1):
                args = {"q": q.detach().clone(), "eps": eps.detach().clone(), "sigma": sig.detach().clone()}
                args[name][i] += sgn * h
                vals.append(float(frame_averaged_energy(eng, frames, args["q"], args["eps"], args["sigma"],
                                                        confs)))
            fd = (vals[0] - vals[1]) / (2 * h)
            assert float(t.grad[i]) == pytest.approx(fd, rel=1e-6, abs=1e-6 * max(1.0, abs(fd)))


def test_autograd_rejects_mismatched_configs():
    syn = synth.tiny_soln()
    eng = OracleEngine(syn.system)
    t = torch.zeros(5, dtype=torch.float64)
    with pytest.raises(ValueError):
        frame_averaged_energy(eng, [(syn.frame(0), syn.cell, 1.0)], t, t, t, [])
