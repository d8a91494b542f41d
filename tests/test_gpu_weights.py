"""Non-unit weights on the device (update_histogram, engproc.F90:758-856):
the self-energy weight (wgtslf = 1, with voffset taken from the first
configuration), the NPT volume correction (volcorrect, engproc.F90:871-887),
frame weights (stat_weight_system) and per-configuration solute weights
(SltWght for flexible solutes), before and after engstore.

The weighted sums are accumulated in another order than the oracle's, and
voffset is the first configuration's self energy (1e-11 relative, like
every energy), so edens / ecorr / eself / engnorm / avslf agree to 1e-12
relative instead of bit for bit; engsmpl is exact.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from ermod_amd import synth
from ermod_amd.api import ErmGPU
from ermod_amd.system import ES_NPT, PT_TEST_FLEX, SLT_REFS_FLEX

pytestmark = pytest.mark.gpu

def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def compare(a, b, keys=("edens", "ecorr", "eself")):
    for k in keys:
        if k in a:
            np.testing.assert_allclose(b[k], a[k], rtol=1e-12, atol=1e-12 * max(1e-300, np.abs(a[k]).max()),
                                       err_msg=k)
    assert b["engsmpl"] == a["engsmpl"]
    assert b["engnorm"] == pytest.approx(a["engnorm"], rel=1e-12)
    assert b["avslf"] == pytest.approx(a["avslf"], rel=1e-12)
    np.testing.assert_allclose(b["minuv"], a["minuv"], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(b["maxuv"], a["maxuv"], rtol=1e-11, atol=1e-9)


def weighted_system(sy, **kw):
    s = sy.system
    return dataclasses.replace(s, wgtslf=1, selfcal=1, self_hist=s.hists[0], **kw)


@pytest.mark.parametrize("estype", [1, ES_NPT])
def test_refs_rigid_self_weight(estype):
    print_case(f"refs rigid wgtslf 1 estype {estype}")
    sy = synth.tiny_refs()
    s = weighted_system(sy, estype=estype)
    m = sy.solute_mol
    o = O.Oracle(s)
    n, seed = 30, 41
    fw = [1.0, 0.6, 1.7]
    with ErmGPU(s, "fp64", max_batch=8) as g:
        for f, w in enumerate(fw):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell, w)
            g.set_frame(pos, sy.cell, w)
            u = g.insert_random(n, seed, first_index=f * n, want_uv=True)
            for i in range(n):
                r, q = O.random_pose(seed, f * n + i)
                want = o.uv_energy(m, O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q))
                np.testing.assert_allclose(u[i], want, rtol=1e-11, atol=1e-9)
                o.update_histogram(m, want)
        a, b = o.acc(), g.fetch()
        assert b["voffset"] == pytest.approx(a["voffset"], rel=1e-11, abs=1e-12)
        compare(a, b)
        o.engstore()
        g.engstore()
        compare(o.acc(), g.fetch())


def test_refs_flexible_structure_weights():
    """slttype 3: host-placed configurations of a flexible solute with
    per-configuration weights (as read from SltWght), self-energy weight."""
    print_case("refs flexible, SltWght weights, wgtslf 1")
    sy = synth.tiny_refs()
    s0 = sy.system
    m = sy.solute_mol
    sluvid = s0.sluvid.copy()
    sluvid[m] = PT_TEST_FLEX
    s = weighted_system(sy, slttype=SLT_REFS_FLEX, sluvid=sluvid)
    rng = np.random.default_rng(3)
    o = O.Oracle(s)
    with ErmGPU(s, "fp64", max_batch=8) as g:
        for f in range(2):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            confs, wts = [], rng.uniform(0.2, 2.0, 20)
            for i in range(20):
                bf = s.bfcoord + rng.normal(scale=0.05, size=s.bfcoord.shape)
                r, q = O.random_pose(9, 100 * f + i)
                confs.append(O.place_solute(bf, s.solute_mass(m), sy.cell, r, q))
            confs = np.array(confs)
            u = g.engconst(solute_pos=confs, weights=wts, want_uv=True)
            for i in range(20):
                want = o.uv_energy(m, confs[i])
                np.testing.assert_allclose(u[i], want, rtol=1e-11, atol=1e-9)
                o.update_histogram(m, want, wts[i])
        a, b = o.acc(), g.fetch()
        assert b["voffset"] == pytest.approx(a["voffset"], rel=1e-11, abs=1e-12)
        compare(a, b)


def test_soln_self_weight():
    print_case("soln wgtslf 1, NPT")
    sy = synth.tiny_soln()
    s = weighted_system(sy, estype=ES_NPT)
    slt = sy.solute_mol
    o = O.Oracle(s)
    with ErmGPU(s, "fp64") as g:
        for f, w in enumerate([1.0, 0.8, 1.3]):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell, w)
            g.set_frame(pos, sy.cell, w)
            want = o.uv_energy(slt)
            got = g.engconst(solutes=[slt], want_uv=True)[0]
            np.testing.assert_allclose(got, want, rtol=1e-11, atol=1e-9)
            o.update_histogram(slt, want)
        a, b = o.acc(), g.fetch()
        compare(a, b, keys=("edens", "eself"))
        np.testing.assert_allclose(b["slnuv"], a["slnuv"], rtol=1e-12)
