"""Random insertions of a flexible test particle on the device
(erm_insert_flexible): body configurations read from a SltConf-style DCD
file through the library's reader, SltWght-style weights, the random and
sphere placement modes; energies against the oracle's placement of the same
bodies and draws (FP64, 1e-9 + 1e-11 relative), weighted accumulators to
1e-12."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from ermod_amd import synth
from ermod_amd.api import ErmGPU, read_dcd
from ermod_amd.system import (INSORG_AGGCEN, INSORG_ORIGIN, INSPOS_RANDOM, INSPOS_SPHERE, INSROT_RANDOM,
                              PT_TEST_FLEX, SLT_REFS_FLEX)

pytestmark = pytest.mark.gpu

TOL64 = dict(rtol=1e-11, atol=1e-9)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


@pytest.mark.parametrize("ins", [(INSORG_ORIGIN, INSPOS_RANDOM, INSROT_RANDOM, 0.0, 0.0),
                                 (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_RANDOM, 1.0, 4.0)])
def test_flexible_insertions_match_oracle(tmp_path, ins):
    print_case(f"flexible test particle, modes {ins[:3]}")
    sy = synth.tiny_refs()
    m = sy.solute_mol
    sluvid = sy.system.sluvid.copy()
    sluvid[m] = PT_TEST_FLEX
    s = dataclasses.replace(sy.system, slttype=SLT_REFS_FLEX, sluvid=sluvid)
    rng = np.random.default_rng(6)
    nconf = 30
    bodies = s.bfcoord[None] + rng.normal(scale=0.08, size=(nconf,) + s.bfcoord.shape)
    synth.write_dcd(tmp_path / "SltConf", bodies, None)
    confs, _ = read_dcd(tmp_path / "SltConf", periodic=False)
    confs = confs.astype(np.float64)
    wts = rng.uniform(0.3, 1.7, nconf)
    o = O.Oracle(s)
    mass = s.solute_mass(m)
    seed = 23
    with ErmGPU(s, "fp64", max_batch=16) as g:
        g.set_insertion(*ins, hostspec=[1])
        for f, (lo, hi) in enumerate([(0, 18), (18, 30)]):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            u = g.insert_flexible(confs[lo:hi], seed, first_index=lo, weights=wts[lo:hi], want_uv=True)
            cen = O.aggregate_center(s, pos, [1])
            for i in range(lo, hi):
                sp = O.insertion_pose(confs[i], mass, sy.cell, ins, seed, i, cen)
                want = o.uv_energy(m, sp)
                np.testing.assert_allclose(u[i - lo], want, **TOL64)
                o.update_histogram(m, want, wts[i])
        a, b = o.acc(), g.fetch()
    np.testing.assert_allclose(b["edens"], a["edens"], rtol=1e-12)
    np.testing.assert_allclose(b["ecorr"], a["ecorr"], rtol=1e-12, atol=1e-12 * np.abs(a["ecorr"]).max())
    assert b["engnorm"] == pytest.approx(wts.sum(), rel=1e-12)
