"""Bare Coulomb (cltype 0, periodic and without a box) and PPPM (cltype 3)
through the device pipeline against the oracle on identical inputs.

Bars, as for PME:
* FP64: per-molecule energies within 1e-9 kcal/mol + 1e-11 relative;
  edens and ecorr bit-exact with unit weights;
* FP32: per-molecule |dE| <= 1e-4 kcal/mol where |E| <= 100 kcal/mol,
  relative 1e-5 above.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from bars import FP32_ABS, fp32_ok, fp32_report
from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU
from ermod_amd.system import EL_COULOMB, EL_PPPM, SYS_NONPERIODIC, SYS_PERIODIC

pytestmark = pytest.mark.gpu

TOL64 = dict(rtol=1e-11, atol=1e-9)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)




def variant(kind, cltype, boxshp=SYS_PERIODIC):
    """A synthetic system with the given electrostatics.  Bare Coulomb with a
    site cutoff leaves charged fragments of the molecules cut at elecut, so
    the lowest bin edge of the larger test-particle system moves down."""
    sy = getattr(synth, kind)()
    s = dataclasses.replace(sy.system, cltype=cltype, boxshp=boxshp)
    if kind == "refs20" and cltype == EL_COULOMB:
        s = dataclasses.replace(s, hists=[dataclasses.replace(h, ecdmin=-60.0) for h in s.hists])
    return sy, s


def scattered_frame(sy, f):
    """A frame with whole molecules moved by a box length in various
    directions: without a box the coordinates must be used as given."""
    pos = sy.frame(f).copy()
    s = sy.system
    rng = np.random.default_rng(100 + f)
    L = np.diag(sy.cell)
    for m in range(s.nmol):
        if rng.random() < 0.3:
            b, e = s.mol_begin[m], s.mol_begin[m + 1]
            pos[b:e] += rng.integers(-1, 2, 3) * L
    return pos


@pytest.mark.parametrize("kind,cltype", [("tiny_refs", EL_COULOMB), ("tiny_refs", EL_PPPM),
                                         ("refs20", EL_COULOMB), ("refs20", EL_PPPM)])
def test_refs_insertions_fp64(kind, cltype):
    print_case(f"{kind} cltype {cltype} insertions fp64")
    sy, s = variant(kind, cltype)
    pos = sy.frame(0)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    n, seed, first = (24, 7, 300) if kind == "tiny_refs" else (12, 8, 900)
    with ErmGPU(s, "fp64", max_batch=16) as g:
        g.set_frame(pos, sy.cell)
        u = g.insert_random(n, seed, first_index=first, want_uv=True)
        b = g.fetch()
    m = sy.solute_mol
    for i in range(n):
        r, q = O.random_pose(seed, first + i)
        sp = O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q)
        want = o.uv_energy(m, sp)
        np.testing.assert_allclose(u[i], want, **TOL64)
        if cltype == EL_COULOMB:
            assert u[i, 0] == 0.0
        o.update_histogram(m, want)
    a = o.acc()
    assert np.array_equal(a["edens"], b["edens"])
    assert np.array_equal(a["ecorr"], b["ecorr"])
    assert a["engnorm"] == b["engnorm"] and a["engsmpl"] == b["engsmpl"]


@pytest.mark.parametrize("cltype,boxshp", [(EL_COULOMB, SYS_PERIODIC), (EL_COULOMB, SYS_NONPERIODIC),
                                           (EL_PPPM, SYS_PERIODIC)])
def test_soln_energies_and_counts_fp64(cltype, boxshp):
    print_case(f"tiny_soln cltype {cltype} boxshp {boxshp} fp64")
    sy, s = variant("tiny_soln", cltype, boxshp)
    o = O.Oracle(s)
    slt = sy.solute_mol
    with ErmGPU(s, "fp64") as g:
        for f in range(3):
            pos = scattered_frame(sy, f) if boxshp == SYS_NONPERIODIC else sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            want = o.uv_energy(slt)
            got = g.engconst(solutes=[slt], want_uv=True)[0]
            np.testing.assert_allclose(got, want, **TOL64)
            o.update_histogram(slt, want)
        b = g.fetch()
    a = o.acc()
    assert np.array_equal(a["edens"], b["edens"])
    np.testing.assert_allclose(b["slnuv"], a["slnuv"], rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("kind,cltype,boxshp", [("refs20", EL_COULOMB, SYS_PERIODIC),
                                                ("refs20", EL_PPPM, SYS_PERIODIC),
                                                ("soln12", EL_COULOMB, SYS_NONPERIODIC)])
def test_fp32_error(kind, cltype, boxshp):
    print_case(f"{kind} cltype {cltype} boxshp {boxshp} fp32 vs oracle")
    sy, s = variant(kind, cltype, boxshp)
    pos = sy.frame(1)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    m = sy.solute_mol
    with ErmGPU(s, "fp32") as g:
        g.set_frame(pos, sy.cell)
        if s.slttype == 1:
            got = g.get_uv_energy(m)
            want = o.uv_energy(m)[None]
        else:
            got = g.insert_random(16, 21, want_uv=True)
            want = []
            for i in range(16):
                r, q = O.random_pose(21, i)
                want.append(o.uv_energy(m, O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q)))
            want = np.array(want)
    fp32_report(f"{kind} cltype {cltype} boxshp {boxshp} vs oracle", got[:, 1:], want[:, 1:])
    assert np.all(np.abs(got[:, 0] - want[:, 0]) < 1e-3)


@pytest.mark.parametrize("path", ["fused", "unfused", "register"])
def test_pppm_every_reciprocal_path(path):
    print_case(f"refs20 PPPM {path} path")
    sy, s = variant("refs20", EL_PPPM)
    pos = sy.frame(2)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    m = sy.solute_mol
    with ErmGPU(s, "fp64", pme_path=path) as g:
        g.set_frame(pos, sy.cell)
        u = g.insert_random(6, 3, want_uv=True)
    for i in range(6):
        r, q = O.random_pose(3, i)
        want = o.uv_energy(m, O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q))
        np.testing.assert_allclose(u[i], want, **TOL64)


def test_pppm_gradient_fp64():
    print_case("tiny_refs PPPM gradient fp64")
    sy, s = variant("tiny_refs", EL_PPPM)
    pos = sy.frame(0)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    m = sy.solute_mol
    confs = []
    for i in range(5):
        r, q = O.random_pose(13, i)
        confs.append(O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q))
    confs = np.array(confs)
    want = [0.0, 0.0, 0.0, 0.0, 0.0]
    for p in confs:
        u, gq, ge, gs = o.solute_gradient(m, p, 0)
        want = [want[0] + 1.0, want[1] + u, want[2] + gq, want[3] + ge, want[4] + gs]
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        got = g.solute_gradient(solute_pos=confs)
    assert got[0] == want[0]
    assert got[1] == pytest.approx(want[1], rel=1e-9, abs=1e-9)
    for a, b in zip(got[2:], want[2:]):
        np.testing.assert_allclose(a, b, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(b).max()))


def test_bare_gradient_is_refused():
    print_case("bare Coulomb gradient refused")
    sy, s = variant("tiny_refs", EL_COULOMB)
    with ErmGPU(s, "fp64") as g:
        g.set_frame(sy.frame(0), sy.cell)
        with pytest.raises(ErmError) as ei:
            g.solute_gradient(n=2, seed=1)
        assert ei.value.code == 11
