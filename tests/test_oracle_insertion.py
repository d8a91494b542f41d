"""Pins of the oracle's insertion modes (instslt, insertion.F90:103-366).

* ORIGIN + RANDOM through orc_insertion_pose is bit for bit the placement
  the rest of the tests use (random_pose + place_solute);
* com_aggregate: numpy's mass-weighted mean of the host molecules' sites;
* SPHERE: the COM lies at lwreg <= |COM - centre| <= upreg, uniformly in
  volume (median of r^3 at the mid value) and isotropically (mean direction
  near zero);
* SLAB: the COM's z offset from the centre lies in [lwreg, upreg] (either
  sign, about half negative, when symmetric) and x, y cover the cell;
* orientation NOCHANGE translates the body rigidly; RANDOM keeps every
  intramolecular distance and the COM.
"""
import numpy as np
import pytest

from ermod_amd import synth
from ermod_amd.system import (INSORG_AGGCEN, INSORG_ORIGIN, INSPOS_RANDOM, INSPOS_SLAB_GENERIC,
                              INSPOS_SLAB_SYMMETRIC, INSPOS_SPHERE, INSROT_NOCHANGE, INSROT_RANDOM)


@pytest.fixture(scope="module")
def body():
    sy = synth.tiny_refs()
    s = sy.system
    m = sy.solute_mol
    return sy, s.bfcoord, s.solute_mass(m)


def com(x, w):
    return (x * w[:, None]).sum(0) / w.sum()


def test_random_mode_is_the_default_placement(orc, body):
    sy, bf, w = body
    for i in range(50):
        r, q = orc.random_pose(17, i)
        want = orc.place_solute(bf, w, sy.cell, r, q)
        got = orc.insertion_pose(bf, w, sy.cell, (INSORG_ORIGIN, INSPOS_RANDOM, INSROT_RANDOM, 0, 0), 17, i)
        assert np.array_equal(got, want)


def test_aggregate_center(orc):
    sy = synth.mixed(20)
    s = sy.system
    pos = sy.frame(0)
    for hosts in ([1], [2], [1, 3, 4]):
        got = orc.aggregate_center(s, pos, hosts)
        sel = orc.host_molecules(s, hosts).astype(bool)
        idx = np.concatenate([np.arange(s.mol_begin[m], s.mol_begin[m + 1]) for m in np.nonzero(sel)[0]])
        want = (pos[idx] * s.sitemass[idx, None]).sum(0) / s.sitemass[idx].sum()
        np.testing.assert_allclose(got, want, rtol=1e-12)


def test_sphere(orc, body):
    sy, bf, w = body
    cen = np.array([3.1, -2.0, 7.5])
    lw, up = 2.0, 6.0
    n = 4000
    shifts = []
    for i in range(n):
        out = orc.insertion_pose(bf, w, sy.cell, (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_NOCHANGE, lw, up), 5, i, cen)
        d = out - (bf - com(bf, w) + cen)
        np.testing.assert_allclose(d, np.broadcast_to(d[0], d.shape), atol=1e-12)  # rigid translation
        shifts.append(com(out, w) - cen)
    sh = np.array(shifts)
    r = np.linalg.norm(sh, axis=1)
    assert r.min() >= lw - 1e-12 and r.max() <= up + 1e-12
    frac = np.mean(r ** 3 <= (lw ** 3 + up ** 3) / 2)
    assert abs(frac - 0.5) < 4 * 0.5 / np.sqrt(n)
    assert np.all(np.abs((sh / r[:, None]).mean(0)) < 4 / np.sqrt(3 * n))


@pytest.mark.parametrize("mode", [INSPOS_SLAB_GENERIC, INSPOS_SLAB_SYMMETRIC])
def test_slab(orc, body, mode):
    sy, bf, w = body
    cen = np.array([1.0, 2.0, 3.0])
    lw, up = (-1.5, 4.0) if mode == INSPOS_SLAB_GENERIC else (1.5, 4.0)
    n = 3000
    L = np.diag(sy.cell)
    dz, dxy = [], []
    for i in range(n):
        out = orc.insertion_pose(bf, w, sy.cell, (INSORG_AGGCEN, mode, INSROT_RANDOM, lw, up), 9, i, cen)
        c = com(out, w) - cen
        dz.append(c[2])
        dxy.append(c[:2] / L[:2])
    dz, dxy = np.array(dz), np.array(dxy)
    if mode == INSPOS_SLAB_GENERIC:
        assert dz.min() >= lw - 1e-9 and dz.max() <= up + 1e-9
    else:
        assert np.abs(dz).min() >= lw - 1e-9 and np.abs(dz).max() <= up + 1e-9
        assert abs(np.mean(dz < 0) - 0.5) < 4 * 0.5 / np.sqrt(n)
    assert dxy.min() >= -1e-9 and dxy.max() < 1 + 1e-9
    assert np.all(np.abs(dxy.mean(0) - 0.5) < 4 * np.sqrt(1 / 12 / n))


def test_random_orientation_is_rigid(orc, body):
    sy, bf, w = body
    cen = np.array([0.5, 0.5, 0.5])
    d0 = np.linalg.norm(bf[:, None] - bf[None], axis=2)
    for i in range(30):
        a = orc.insertion_pose(bf, w, sy.cell, (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_NOCHANGE, 1.0, 3.0), 2, i, cen)
        b = orc.insertion_pose(bf, w, sy.cell, (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_RANDOM, 1.0, 3.0), 2, i, cen)
        np.testing.assert_allclose(com(a, w), com(b, w), atol=1e-12)
        np.testing.assert_allclose(np.linalg.norm(b[:, None] - b[None], axis=2), d0, atol=1e-12)
