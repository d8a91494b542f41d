"""Reference-structure insertions on the device (erm_set_refstructure +
insorigin REFSTR with RMSD or GAUSS positions) against the oracle's reffit
and placement of the same draws: FP64 energies to 1e-9 + 1e-11 relative,
unit-weight counts bit-exact (RMSD), GAUSS-weighted accumulators to 1e-12,
over two frames (the fit follows each frame)."""
import numpy as np
import pytest

import oracle as O
from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU
from ermod_amd.system import INSORG_REFSTR, INSPOS_GAUSS, INSPOS_RMSD, INSROT_RANDOM

pytestmark = pytest.mark.gpu

TOL64 = dict(rtol=1e-11, atol=1e-9)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def reference(sy):
    s = sy.system
    m = sy.solute_mol
    pos = sy.frame(0)
    host = np.arange(0, 30, dtype=np.int32)
    rng = np.random.default_rng(3)
    host_ref = pos[host] + rng.normal(scale=0.2, size=(30, 3))
    host_w = np.where(s.sitemass[host] > 0.95 * 4.0026, 1.0, 0.0)
    slt_ref = s.bfcoord - s.bfcoord.mean(0) + host_ref.mean(0) + np.array([2.0, 0.0, 0.0])
    slt_w = np.ones(s.nsite(m))
    return host, host_ref, host_w, slt_ref, slt_w


@pytest.mark.parametrize("mode,lw,up", [(INSPOS_RMSD, 0.5, 3.0), (INSPOS_GAUSS, 0.0, 4.0)])
def test_refstr_insertions_match_oracle(mode, lw, up):
    print_case(f"reference-structure insertion mode {mode}")
    sy = synth.tiny_refs()
    s = sy.system
    m = sy.solute_mol
    mass = s.solute_mass(m)
    host, host_ref, host_w, slt_ref, slt_w = reference(sy)
    ins = (INSORG_REFSTR, mode, INSROT_RANDOM, lw, up)
    g = O.gauss_constants(sy.cell, up)
    o = O.Oracle(s)
    n, seed = 24, 19
    with ErmGPU(s, "fp64", max_batch=16) as dev:
        dev.set_refstructure(host, host_ref, host_w, slt_ref, slt_w)
        dev.set_insertion(*ins)
        for f in (1, 2):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            dev.set_frame(pos, sy.cell)
            fitted, best = O.reffit(pos, host, host_ref, host_w, slt_ref, slt_w, s.bfcoord)
            u = dev.insert_random(n, seed, first_index=f * n, want_uv=True)
            for i in range(n):
                sp, wt = O.refstr_pose(mass, sy.cell, ins, fitted, best, slt_w, g, seed, f * n + i)
                want = o.uv_energy(m, sp)
                np.testing.assert_allclose(u[i], want, **TOL64)
                o.update_histogram(m, want, wt)
        a, b = o.acc(), dev.fetch()
    if mode == INSPOS_RMSD:
        assert np.array_equal(a["edens"], b["edens"])
        assert np.array_equal(a["ecorr"], b["ecorr"])
    else:
        np.testing.assert_allclose(b["edens"], a["edens"], rtol=1e-12)
        np.testing.assert_allclose(b["ecorr"], a["ecorr"], rtol=1e-12, atol=1e-12 * np.abs(a["ecorr"]).max())
    assert b["engnorm"] == pytest.approx(a["engnorm"], rel=1e-12)


def test_refstr_argument_checks():
    print_case("reference-structure argument checks")
    sy = synth.tiny_refs()
    host, host_ref, host_w, slt_ref, slt_w = reference(sy)
    with ErmGPU(sy.system, "fp64") as dev:
        with pytest.raises(ErmError):          # no reference structure yet
            dev.set_insertion(INSORG_REFSTR, INSPOS_RMSD, 0, 0.0, 1.0)
        dev.set_refstructure(host, host_ref, host_w, slt_ref, slt_w)
        with pytest.raises(ErmError):          # lwreg > upreg
            dev.set_insertion(INSORG_REFSTR, INSPOS_RMSD, 0, 2.0, 1.0)
        with pytest.raises(ErmError):          # RMSD needs REFSTR
            dev.set_insertion(0, INSPOS_RMSD, 0, 0.0, 1.0)
        with pytest.raises(ErmError):
            dev.set_refstructure(host, host_ref, np.zeros_like(host_w), slt_ref, slt_w)
