"""Pins of the oracle's intermediate-state accumulation (solution systems,
extensions/Intermediate/engproc.F90:996-1009, 348-352, 656):

* with no empty reference bin every configuration is intermediate, so the
  accumulators equal edens / slnuv / engnorm;
* with a mask, the configurations kept are exactly those whose
  instantaneous histogram touches no masked bin (brute force over the
  returned histograms);
* engclear clears edintm, uvintm and norm_intm but not soln_intm and
  soln_total; engstore divides edintm by norm_intm.
"""
import numpy as np

from ermod_amd import synth


def run(orc, mask, frames=(0, 1, 2, 3)):
    sy = synth.soln12(nwater=400, L=23.0, K=24)
    s = sy.system
    o = orc.Oracle(s)
    o.set_intermediate(mask(o) if callable(mask) else mask)
    hs, us = [], []
    for f in frames:
        o.set_frame(sy.frame(f), sy.cell)
        u = o.uv_energy(sy.solute_mol)
        hs.append(o.update_histogram(sy.solute_mol, u))
        us.append(u)
    return o, np.array(hs), np.array(us)


def test_no_empty_bin_keeps_everything(orc):
    o, _, _ = run(orc, lambda o: np.zeros(o.ermax, np.int32))
    a = o.acc()
    ed, uv, norm, sint, stot = o.intermediate()
    assert np.array_equal(ed, a["edens"])
    np.testing.assert_array_equal(uv[1:], a["slnuv"][1:])
    assert norm == a["engnorm"] == sint == stot == 4.0


def test_mask_selects_configurations(orc):
    o0, hs, us = run(orc, lambda o: np.zeros(o.ermax, np.int32))
    # empty-in-reference bins: the highest occupied bin of the first frame
    mask = np.zeros(o0.ermax, np.int32)
    mask[np.nonzero(hs[0])[0].max()] = 1
    o, hs2, _ = run(orc, mask)
    assert np.array_equal(hs, hs2)
    keep = ~np.any((hs > 0) & (mask[None, :] > 0), axis=1)
    assert 0 < keep.sum() < len(keep)
    ed, uv, norm, sint, stot = o.intermediate()
    assert np.array_equal(ed, hs[keep].sum(0).astype(float))
    np.testing.assert_allclose(uv[1], us[keep, 1:].sum(), rtol=1e-12)
    assert (norm, sint, stot) == (keep.sum(), keep.sum(), len(keep))
    o.engclear()
    ed, uv, norm, sint, stot = o.intermediate()
    assert not ed.any() and not uv.any() and norm == 0.0
    assert (sint, stot) == (keep.sum(), len(keep))


def test_engstore_normalises(orc):
    o, hs, _ = run(orc, lambda o: np.zeros(o.ermax, np.int32))
    o.engstore()
    ed, _, norm, _, _ = o.intermediate()
    np.testing.assert_allclose(ed, hs.sum(0) / norm, rtol=1e-15)
