"""Pins of the oracle's bare-Coulomb and PPPM branches (cltype 0 and 3).

* calc_gfb_pppm: the coefficients are those of the alias sum of the squared
  charge-assignment function, sum_m sinc(u + m)^(2P) = sum_l b_l sin^2l(pi u)
  (Hockney & Eastwood), checked against the sum itself;
* the PPPM Green's function: the optimal influence function evaluated
  independently with a wider alias range, on a grid coarse enough that the
  reference's alias range is not empty;
* PPPM energies: a direct Ewald k-sum and the NaCl Madelung constant;
* realcal_bare: closed forms of Coulomb and LJ for isolated pairs, the
  minimum image by enumeration of the 27 neighbouring images, the Coulomb
  cutoff when periodic and none otherwise, and agreement with the screened
  real-space sum at zero screening (an independent LJ-switch code path).
"""
import dataclasses

import numpy as np
import pytest

from ermod_amd.system import CHARGE_SCALE, EL_COULOMB, EL_PPPM, SYS_NONPERIODIC

from test_oracle_pins import ewald_recip_pair, nacl_system, point_charge_system


@pytest.mark.parametrize("order", [2, 4, 6, 8])
def test_pppm_gfb_is_alias_sum(orc, order):
    b = orc.gfb_pppm(order)
    m = np.arange(-20000, 20001)
    for u in (0.07, 0.23, 0.37, 0.5):
        t = np.pi * (u + m)
        brute = np.sum((np.sin(t) / t) ** (2 * order))
        x = np.sin(np.pi * u) ** 2
        assert abs(sum(b[l] * x ** l for l in range(order)) - brute) < 1e-13


def test_pppm_green_vs_optimal_influence(orc):
    """The optimal influence function written independently: numerator over
    the alias range the reference chooses per axis, int(L_d alpha / (pi K_d)
    (-ln 1e-7)^(1/4)) (here 0 along x, 1 along y and z), denominator from a
    wide alias sum instead of the gf_b polynomial."""
    K = (8, 6, 6)
    Ls = np.array([13.0, 11.0, 11.0])
    cell = np.diag(Ls)
    p, alpha = 4, 0.9
    got = orc.pppm_greenfunc(K, p, cell, alpha)
    mmax = [int(Ls[d] * alpha / (np.pi * K[d]) * (-np.log(1e-7)) ** 0.25) for d in range(3)]
    assert mmax == [0, 1, 1]
    inv = np.linalg.inv(cell)
    vol = np.linalg.det(cell)
    fold = [np.where(np.arange(k) <= k // 2, np.arange(k), np.arange(k) - k).astype(float) for k in K]
    H = K[0] // 2 + 1
    want = np.zeros_like(got)
    A = np.stack(np.meshgrid(*[np.arange(-mm, mm + 1) for mm in mmax], indexing="ij"), -1).reshape(-1, 3)
    wide = np.arange(-4000, 4001)
    for i3 in range(K[2]):
        for i2 in range(K[1]):
            for i1 in range(H):
                if i1 == 0 and i2 == 0 and i3 == 0:
                    continue
                m = np.array([fold[0][i1], fold[1][i2], fold[2][i3]])
                k = 2 * np.pi * (inv.T @ m)
                mk = m[None, :] + A * np.array(K)[None, :]
                km = 2 * np.pi * mk @ inv
                U2 = np.prod(np.sinc(mk / np.array(K)[None, :]) ** (2 * p), axis=1)
                gam = np.exp(-(km ** 2).sum(1) / (4 * alpha ** 2))
                num = np.sum((km @ k) / (km ** 2).sum(1) * U2 * gam)
                den = np.prod([np.sum(np.sinc(m[d] / K[d] + wide) ** (2 * p)) for d in range(3)]) ** 2
                want[i3, i2, i1] = 4 * np.pi * num / ((k @ k) * den * vol)
    np.testing.assert_allclose(got, want, rtol=1e-10, atol=1e-14 * np.abs(want).max())


@pytest.mark.parametrize("K,p,tol", [(48, 8, 1e-9), (32, 6, 2e-7)])
def test_pppm_recip_vs_direct_ewald_sum(orc, K, p, tol):
    L, alpha = 12.0, 0.45
    rng = np.random.default_rng(11)
    pts = [[*rng.uniform(0, L, 3), rng.uniform(-1, 1), 3.0, 0.0] for _ in range(6)]
    s, pos = point_charge_system(pts, L, K, p, alpha, 5.9)
    s = dataclasses.replace(s, cltype=EL_PPPM)
    o = orc.Oracle(s)
    o.set_frame(pos, np.diag([L, L, L]))
    erec, self_e = o.recpcal(0)
    q = s.charge
    for k in range(1, 6):
        want = q[0] * q[k] * ewald_recip_pair(pos[0] - pos[k], L, alpha)
        assert abs(erec[k] - want) < tol * max(1.0, abs(want)), (k, erec[k], want)
    want_self = 0.5 * q[0] ** 2 * ewald_recip_pair(np.zeros(3), L, alpha)
    assert abs(self_e - want_self) < tol * abs(want_self)


def test_pppm_madelung(orc):
    (s, pos), L = nacl_system()
    s = dataclasses.replace(s, cltype=EL_PPPM)
    o = orc.Oracle(s)
    o.set_frame(pos, np.diag([L, L, L]))
    u = o.uv_energy(0)
    site = u[1:].sum() + 2.0 * u[0]
    want = -1.747564594633 * CHARGE_SCALE ** 2 / 2.5
    assert abs(site - want) < 1e-8 * abs(want), (site, want)


def test_pppm_and_nonperiodic_argument_checks(orc):
    pts = [[1.0, 1.0, 1.0, 0.5, 3.0, 0.1], [4.0, 1.0, 1.0, -0.5, 3.0, 0.1]]
    s, _ = point_charge_system(pts, 20.0, 16, 5, 0.35, 9.0)
    with pytest.raises(orc.OracleError):
        orc.Oracle(dataclasses.replace(s, cltype=EL_PPPM))           # odd order
    with pytest.raises(orc.OracleError):
        orc.Oracle(dataclasses.replace(s, cltype=1))                 # Ewald
    with pytest.raises(orc.OracleError):
        orc.Oracle(dataclasses.replace(s, boxshp=SYS_NONPERIODIC))   # non-periodic PME


def test_bare_pair_closed_forms(orc):
    """Isolated pairs without a box: q q / r at any distance, LJ zero at
    sigma and -eps at 2^(1/6) sigma, nothing beyond upljcut."""
    eps_a, eps_b, sg_a, sg_b = 0.15, 0.12, 3.2, 3.4
    eps, sig = np.sqrt(eps_a * eps_b), 0.5 * (sg_a + sg_b)
    for r, want_lj in ((sig, 0.0), (2 ** (1 / 6) * sig, -eps), (25.0, 0.0), (400.0, 0.0)):
        pts = [[0.3, -0.2, 0.1, 0.4, sg_a, eps_a], [0.3 + r, -0.2, 0.1, -0.7, sg_b, eps_b]]
        s, pos = point_charge_system(pts, 30.0, 16, 4, 0.35, 12.0)
        s = dataclasses.replace(s, cltype=EL_COULOMB, boxshp=SYS_NONPERIODIC)
        o = orc.Oracle(s)
        o.set_frame(pos, np.zeros((3, 3)))
        u = o.uv_energy(0)
        want = CHARGE_SCALE ** 2 * 0.4 * -0.7 / r + want_lj
        assert u[0] == 0.0
        assert abs(u[2] - want) < 1e-12 * max(1.0, abs(want)), (r, u[2], want)


def test_bare_periodic_minimum_image_and_cutoff(orc):
    """Periodic bare Coulomb: each site pair at its nearest image among the
    27 neighbouring cells, Coulomb dropped at r >= elecut, plain LJ inside
    lwljcut (every pair here has r < lwljcut or r > upljcut)."""
    rng = np.random.default_rng(5)
    L = np.array([16.0, 19.0, 17.0])
    rc = 7.5
    pts = [[*rng.uniform(-L, 2 * L), rng.uniform(-1, 1), rng.uniform(2.5, 3.5), rng.uniform(0.05, 0.2)]
           for _ in range(60)]
    s, pos = point_charge_system(pts, 16.0, 16, 4, 0.35, rc, lw=rc)
    s = dataclasses.replace(s, cltype=EL_COULOMB, lwljcut=rc, upljcut=rc)
    o = orc.Oracle(s)
    o.set_frame(pos, np.diag(L))
    got = o.uv_energy(0)
    arr = np.array(pts)
    shifts = np.stack(np.meshgrid(*[np.arange(-2, 3)] * 3, indexing="ij"), -1).reshape(-1, 3) * L
    x0 = np.mod(arr[0, :3], L)
    # uvengy(1 + k) belongs to molecule k (every molecule is physical);
    # molecule 0 is the solute itself
    assert got[0] == 0.0 and got[1] == 0.0
    for k in range(1, len(arr)):
        d = np.mod(arr[k, :3], L) - x0
        r = np.min(np.linalg.norm(d[None, :] + shifts, axis=1))
        e = 0.0
        if r < rc:
            e += CHARGE_SCALE ** 2 * arr[0, 3] * arr[k, 3] / r
            eps = np.sqrt(arr[0, 5] * arr[k, 5])
            sig = 0.5 * (arr[0, 4] + arr[k, 4])
            e += 4 * eps * ((sig / r) ** 12 - (sig / r) ** 6)
        assert abs(got[1 + k] - e) < 1e-11 * max(1.0, abs(e)), (k, got[1 + k], e)


@pytest.mark.parametrize("ljswitch", [0, 1, 2, 3])
def test_bare_equals_unscreened_real_space(orc, ljswitch):
    """Periodic bare Coulomb is the screened real-space sum at alpha = 0: the
    two restatements share no code for the LJ switches."""
    from ermod_amd import synth
    sy = synth.soln12(nwater=700, L=28.0)
    s = dataclasses.replace(sy.system, ljswitch=ljswitch, screen=0.0)
    ob = orc.Oracle(dataclasses.replace(s, cltype=EL_COULOMB))
    ob.set_frame(sy.frame(1), sy.cell)
    bare = ob.realcal_bare(sy.solute_mol)
    uv = ob.uv_energy(sy.solute_mol)
    assert uv[0] == 0.0
    np.testing.assert_array_equal(uv[1:], bare)
    o = orc.Oracle(dataclasses.replace(s, screen=1e-300))
    o.set_frame(sy.frame(1), sy.cell)
    brute = o.realcal(sy.solute_mol, brute=True)
    np.testing.assert_allclose(bare, brute, rtol=1e-12, atol=1e-12)
    assert np.count_nonzero(bare) > 100
