"""Pins of the oracle's reference-structure insertion (insorigin REFSTR):

* fit_a_rotate_b: the weighted superposition equals the Kabsch (SVD)
  solution computed with numpy, and an exactly rotated and shifted copy is
  brought back onto its original;
* rmsd_nofit: its closed form;
* RMSD mode: every accepted pose has lwreg <= RMSD <= upreg against the
  best-fit reference solute and a COM within upreg + 3 * (largest COM
  distance) of the fitted solute's COM along each axis;
* GAUSS mode: with a window that accepts the first attempt, the weight is
  1 / (0.5 / z0 + 0.5 exp(-|s|^2) / z1) with s recovered from the COM shift.
"""
import numpy as np
import pytest

from ermod_amd import synth
from ermod_amd.system import INSORG_REFSTR, INSPOS_GAUSS, INSPOS_RMSD, INSROT_NOCHANGE, INSROT_RANDOM


def rot(rng):
    q = rng.normal(size=4)
    q /= np.linalg.norm(q)
    w, x, y, z = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def kabsch(ref, mov, w, b):
    cr = (ref * w[:, None]).sum(0) / w.sum()
    cm = (mov * w[:, None]).sum(0) / w.sum()
    H = ((mov - cm) * w[:, None]).T @ (ref - cr)
    U, S, Vt = np.linalg.svd(H)
    d = np.sign(np.linalg.det(Vt.T @ U.T))
    R = Vt.T @ np.diag([1, 1, d]) @ U.T
    return (b - cm) @ R.T + cr


def test_fit_is_kabsch(orc):
    rng = np.random.default_rng(1)
    for _ in range(10):
        a = rng.normal(scale=3.0, size=(12, 3))
        w = rng.uniform(0.2, 2.0, 12)
        w[rng.integers(0, 12, 3)] = 0.0
        ref = a @ rot(rng).T + rng.normal(size=3) * 5 + rng.normal(scale=0.3, size=(12, 3))
        b = rng.normal(scale=2.0, size=(7, 3))
        np.testing.assert_allclose(orc.fit_a_rotate_b(ref, a, w, b), kabsch(ref, a, w, b), atol=1e-10)
        moved = a @ rot(rng).T + np.array([1.0, -2.0, 3.5])
        np.testing.assert_allclose(orc.fit_a_rotate_b(a, moved, w, moved), a, atol=1e-11)


def test_rmsd_nofit(orc):
    rng = np.random.default_rng(2)
    a, b, w = rng.normal(size=(9, 3)), rng.normal(size=(9, 3)), rng.uniform(0, 1, 9)
    want = np.sqrt((w * ((b - a) ** 2).sum(1)).sum() / w.sum())
    assert orc.rmsd_nofit(a, b, w) == pytest.approx(want, rel=1e-14)


def setup(orc):
    sy = synth.tiny_refs()
    s = sy.system
    m = sy.solute_mol
    pos = sy.frame(0)
    host = np.arange(0, 30)
    rng = np.random.default_rng(3)
    host_ref = pos[host] + rng.normal(scale=0.2, size=(30, 3))
    host_w = np.where(s.sitemass[host] > 0.95 * 4.0026, 1.0, 0.0)
    slt_ref = s.bfcoord - s.bfcoord.mean(0) + host_ref.mean(0) + np.array([2.0, 0.0, 0.0])
    slt_w = np.ones(s.nsite(m))
    fitted, best = orc.reffit(sy.frame(1), host, host_ref, host_w, slt_ref, slt_w, s.bfcoord)
    return sy, s, m, fitted, best, slt_w


def test_rmsd_mode(orc):
    sy, s, m, fitted, best, w = setup(orc)
    mass = s.solute_mass(m)
    lw, up = 0.5, 2.5
    cf = (fitted * mass[:, None]).sum(0) / mass.sum()
    maxdis = np.sqrt(((fitted - cf) ** 2).sum(1).max())
    rmsds = []
    for i in range(300):
        out, wt = orc.refstr_pose(mass, sy.cell, (INSORG_REFSTR, INSPOS_RMSD, INSROT_RANDOM, lw, up), fitted, best,
                                  w, np.zeros(5), 7, i)
        r = orc.rmsd_nofit(best, out, w)
        assert lw <= r <= up and wt == 1.0
        c = (out * mass[:, None]).sum(0) / mass.sum()
        assert np.all(np.abs(c - cf) <= up + 3 * maxdis + 1e-9)
        rmsds.append(r)
    assert np.ptp(rmsds) > 0.5


def test_gauss_mode_weight(orc):
    sy, s, m, fitted, best, w = setup(orc)
    mass = s.solute_mass(m)
    up = 2.0
    g = orc.gauss_constants(sy.cell, up)
    inv = np.linalg.inv(sy.cell)
    for i in range(200):
        out, wt = orc.refstr_pose(mass, sy.cell, (INSORG_REFSTR, INSPOS_GAUSS, INSROT_NOCHANGE, 0.0, 1e9), fitted,
                                  best, w, g, 5, i)
        shift = out[0] - fitted[0]
        np.testing.assert_allclose(out - fitted, np.broadcast_to(shift, out.shape), atol=1e-11)
        sc = (inv @ shift) * 2 * g[:3]
        want = 1.0 / (0.5 / g[3] + 0.5 * np.exp(-(sc ** 2).sum()) / g[4])
        assert wt == pytest.approx(want, rel=1e-10)
