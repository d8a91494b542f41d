"""Parity of the adjoint kernels (erm_solute_gradient through the C-ABI) with
the oracle's FP64 reference gradient, and with central finite differences.

Bars:
* FP64: energy and every gradient component within 1e-9 relative (+1e-9
  times the component's scale) of the oracle, which sums the real-space
  pairs by brute force and the reciprocal part through its own FFT;
* FP32 (pair and PME arithmetic in FP32, per-site sums in FP64): every
  gradient component within 1e-4 of the largest magnitude of its kind
  (q, eps or sigma), the energy within 1e-4 kcal/mol per solvent molecule
  within the cutoff sphere (bounded here by 1e-4 * slvmax);
* finite differences of the device energy (h = 1e-4) agree with the device
  gradient to 1e-6 relative.
"""
import numpy as np
import pytest
import torch

import oracle as O
from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU
from ermod_amd.autograd import frame_averaged_energy
from ermod_amd.system import CHARGE_SCALE, lj_tables, site_lj
import dataclasses

pytestmark = pytest.mark.gpu


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def calm_poses(sy, o, n, seed, limit=100.0):
    """First n random poses (seed) whose |u| < limit, as (indices, coordinates)."""
    s = sy.system
    m = sy.solute_mol
    idx, out = [], []
    i = 0
    while len(out) < n and i < 20000:
        r, q = O.random_pose(seed, i)
        p = O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q)
        if abs(np.sum(o.uv_energy(m, p)[1:])) < limit:
            idx.append(i)
            out.append(p)
        i += 1
    return idx, np.array(out)


def with_switch(sy, ljswitch, cmbrule):
    s = sy.system
    ep, sg = site_lj(s)
    ljtype, epsm, sig2 = lj_tables(ep, sg, cmbrule)
    return dataclasses.replace(s, ljswitch=ljswitch, ljtype=ljtype, lj_eps=epsm, lj_sig2=sig2)


def oracle_sums(o, slt, confs, cmbrule, **kw):
    acc = [0.0, 0.0, 0.0, 0.0, 0.0]
    for p in confs:
        u, gq, ge, gs = o.solute_gradient(slt, p, cmbrule, **kw)
        acc = [acc[0] + 1.0, acc[1] + u, acc[2] + gq, acc[3] + ge, acc[4] + gs]
    return acc


def assert_close64(got, want):
    assert got[0] == want[0]
    assert got[1] == pytest.approx(want[1], rel=1e-9, abs=1e-9)
    for g, w in zip(got[2:], want[2:]):
        np.testing.assert_allclose(g, w, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(w).max()))


def assert_close32(got, want, slvmax, nconf):
    assert abs(got[1] - want[1]) <= 1e-4 * slvmax * nconf
    for g, w in zip(got[2:], want[2:]):
        err = np.abs(g - w).max()
        assert err <= 1e-4 * max(1.0, np.abs(w).max()), err


@pytest.mark.parametrize("kind,ljswitch,cmbrule", [
    ("refs", 0, 0), ("refs", 1, 1), ("refs", 2, 0), ("soln", 3, 1), ("soln", 0, 0)])
def test_gradient_tiny_fp64(kind, ljswitch, cmbrule):
    print_case(f"gradient tiny {kind} switch {ljswitch} cmbrule {cmbrule}")
    sy = synth.tiny_refs() if kind == "refs" else synth.tiny_soln()
    s = with_switch(sy, ljswitch, cmbrule)
    pos = sy.frame(0)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    slt = sy.solute_mol
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        if kind == "refs":
            _, confs = calm_poses(sy, o, 6, 3, limit=1e12)
            got = g.solute_gradient(solute_pos=confs, cmbrule=cmbrule)
        else:
            confs = [None]
            got = g.solute_gradient(solutes=[slt], cmbrule=cmbrule)
    want = oracle_sums(o, slt, confs, cmbrule)
    assert_close64(got, want)


def test_gradient_energy_is_uv_sum():
    print_case("gradient energy vs get_uv_energy")
    sy = synth.tiny_refs()
    pos = sy.frame(1)
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    _, confs = calm_poses(sy, o, 4, 9)
    with ErmGPU(sy.system, "fp64") as g:
        g.set_frame(pos, sy.cell)
        for p in confs:
            uv = g.get_uv_energy(sy.solute_mol, p[None])
            a = g.solute_gradient(solute_pos=p[None])
            assert a[1] == pytest.approx(float(np.sum(uv[0, 1:])), rel=1e-10, abs=1e-9)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gradient_random_poses(prec):
    print_case(f"gradient random poses {prec}")
    sy = synth.tiny_refs()
    pos = sy.frame(2)
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    n, seed, first = 40, 77, 1000
    confs = []
    for i in range(n):
        r, q = O.random_pose(seed, first + i)
        confs.append(O.place_solute(sy.system.bfcoord, sy.system.solute_mass(sy.solute_mol), sy.cell, r, q))
    want = oracle_sums(o, sy.solute_mol, confs, 0)
    with ErmGPU(sy.system, prec) as g:
        g.set_frame(pos, sy.cell)
        if prec == "fp64":
            got = g.solute_gradient(n=n, seed=seed, first_index=first)
            assert_close64(got, want)
        else:
            # FP32 on calm poses only: overlaps amplify distance rounding
            idx, calm = calm_poses(sy, o, 20, seed)
            want = oracle_sums(o, sy.solute_mol, calm, 0)
            got = g.solute_gradient(solute_pos=calm)
            assert_close32(got, want, o.slvmax, len(calm))


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_gradient_soln30(prec):
    """The differentiable-mode system: 30-site solute in 4000 waters, K = 48."""
    print_case(f"gradient soln30 {prec}")
    sy = synth.soln30()
    s = sy.system
    slt = sy.solute_mol
    o = O.Oracle(s)
    with ErmGPU(s, prec) as g:
        for f in range(2):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            want = oracle_sums(o, slt, [None], 0)
            got = g.solute_gradient(solutes=[slt])
            if prec == "fp64":
                assert_close64(got, want)
            else:
                assert_close32(got, want, o.slvmax, 1)


def test_gradient_finite_differences_fp64():
    print_case("gradient finite differences fp64")
    sy = synth.tiny_soln()
    s = sy.system
    slt = sy.solute_mol
    b, e = s.mol_begin[slt], s.mol_begin[slt + 1]
    ep, sg = site_lj(s)
    base = {"q": s.charge[b:e].copy(), "eps": ep[b:e].copy(), "sigma": sg[b:e].copy()}
    pos = sy.frame(0)
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        a = g.solute_gradient(solutes=[slt], **base)
        h = 1e-4
        for k, name in enumerate(("q", "eps", "sigma")):
            for i in range(e - b):
                vals = []
                for sgn in (1, -1):
                    p = {kk: v.copy() for kk, v in base.items()}
                    p[name][i] += sgn * h
                    vals.append(g.solute_gradient(solutes=[slt], **p)[1])
                fd = (vals[0] - vals[1]) / (2 * h)
                assert a[2 + k][i] == pytest.approx(fd, rel=1e-6, abs=1e-6 * max(1.0, abs(fd)))


def test_autograd_function_on_device():
    """frame_averaged_energy over device-resident frames: backward equals the
    oracle's weighted average gradient (charges in e)."""
    print_case("autograd function")
    sy = synth.tiny_soln()
    s = sy.system
    slt = sy.solute_mol
    b, e = s.mol_begin[slt], s.mol_begin[slt + 1]
    ep, sg = site_lj(s)
    dev = torch.device("cuda:0")
    q = torch.tensor(s.charge[b:e] / CHARGE_SCALE, dtype=torch.float64, device=dev, requires_grad=True)
    eps = torch.tensor(ep[b:e], dtype=torch.float64, device=dev, requires_grad=True)
    sig = torch.tensor(sg[b:e], dtype=torch.float64, device=dev, requires_grad=True)
    ws = [1.0, 0.25, 3.0]
    frames = [(torch.tensor(sy.frame(i), dtype=torch.float64, device=dev).contiguous(), sy.cell, w)
              for i, w in enumerate(ws)]
    o = O.Oracle(s)
    acc = np.zeros(1 + 3 * (e - b))
    for i, w in enumerate(ws):
        o.set_frame(sy.frame(i), sy.cell, w)
        u, gq, ge, gs = o.solute_gradient(slt, None, 0)
        acc += w * np.concatenate([[u], gq * CHARGE_SCALE, ge, gs])
    acc /= sum(ws)
    with ErmGPU(s, "fp64") as g:
        u = frame_averaged_energy(g, frames, q, eps, sig, [{"solutes": [slt]}] * 3)
        u.backward()
        torch.cuda.synchronize()
    assert float(u.detach()) == pytest.approx(acc[0], rel=1e-9)
    n = e - b
    np.testing.assert_allclose(q.grad.cpu().numpy(), acc[1:1 + n], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(eps.grad.cpu().numpy(), acc[1 + n:1 + 2 * n], rtol=1e-9, atol=1e-9)
    np.testing.assert_allclose(sig.grad.cpu().numpy(), acc[1 + 2 * n:], rtol=1e-9, atol=1e-9)


def test_gradient_errors():
    print_case("gradient error codes")
    sy = synth.tiny_refs()
    with ErmGPU(sy.system, "fp64") as g:
        with pytest.raises(ErmError):
            g.solute_gradient(n=1)  # no frame
        g.set_frame(sy.frame(0), sy.cell)
        with pytest.raises(ErmError):
            g.solute_gradient(n=1, cmbrule=2)
