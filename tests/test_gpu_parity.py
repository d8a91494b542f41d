"""Parity of the device pipeline (libermgpu.so through its C-ABI) with the
oracle on identical seeded inputs.

Bars:
* FP64: per-molecule energies within 1e-9 kcal/mol + 1e-11 relative of the
  oracle (the two differ only in summation order, FFT algorithm and the
  libm/ocml erf); histogram counts, edens and ecorr bit-exact with unit
  weights;
* FP32 (pair terms and PME in FP32, per-molecule sums in FP64): per-molecule
  |dE| <= 1e-4 kcal/mol where |E| <= 1e4 kcal/mol (the last finite bin
  edge), relative 1e-7 above (overlapping insertions in the overflow bin,
  whose r^-12 repulsion amplifies the FP32 distance rounding); the bin-flip
  fraction is reported (tests/bars.py).
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from bars import FP32_ABS, FP32_EMAX, fp32_ok, fp32_report
from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU

pytestmark = pytest.mark.gpu

TOL64 = dict(rtol=1e-11, atol=1e-9)


def poses(sy, n, seed):
    s = sy.system
    m = sy.solute_mol
    out = []
    for i in range(n):
        r, q = O.random_pose(seed, i)
        out.append(O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q))
    return np.array(out)


@pytest.fixture(scope="module")
def tiny():
    sy = synth.tiny_refs()
    return sy, sy.frame(0)


@pytest.fixture(scope="module")
def refs():
    sy = synth.refs20()
    return sy, sy.frame(0)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


# ------------------------------------------------------------- stage parity
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_realcal_tiny(tiny, prec):
    print_case(f"realcal tiny {prec}")
    sy, pos = tiny
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    with ErmGPU(sy.system, prec) as g:
        g.set_frame(pos, sy.cell)
        for sp in poses(sy, 6, 1):
            want = o.realcal(sy.solute_mol, sp)
            got = g.realcal_proc(sy.solute_mol, sp)
            if prec == "fp64":
                np.testing.assert_allclose(got, want, **TOL64)
            else:
                ok, err = fp32_ok(got, want)
                assert ok.all(), err.max()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_recpcal_tiny(tiny, prec):
    print_case(f"recpcal tiny {prec}")
    sy, pos = tiny
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    with ErmGPU(sy.system, prec) as g:
        g.set_frame(pos, sy.cell)
        for sp in poses(sy, 4, 2):
            want, wself = o.recpcal(sy.solute_mol, sp)
            got, gself = g.recpcal(sy.solute_mol, sp)
            if prec == "fp64":
                np.testing.assert_allclose(got, want, **TOL64)
                assert abs(gself - wself) <= 1e-9 + 1e-11 * abs(wself)
            else:
                ok, err = fp32_ok(got, want)
                assert ok.all(), err.max()
                assert abs(gself - wself) <= 1e-3


def test_realcal_self(tiny):
    print_case("realcal_self tiny")
    sy, pos = tiny
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    with ErmGPU(sy.system, "fp64") as g:
        g.set_frame(pos, sy.cell)
        for sp in poses(sy, 3, 3):
            assert abs(g.realcal_self(sy.solute_mol, sp) - o.realcal_self(sy.solute_mol, sp)) < 1e-10


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_uv_energy_refs20(refs, prec):
    """4000 waters, 49.3 A, 20-site test particle, PME 48^3."""
    print_case(f"uv_energy refs20 {prec}")
    sy, pos = refs
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    P = poses(sy, 5, 4)
    with ErmGPU(sy.system, prec) as g:
        g.set_frame(pos, sy.cell)
        got = g.get_uv_energy(sy.solute_mol, P)
    for i, sp in enumerate(P):
        want = o.uv_energy(sy.solute_mol, sp)
        if prec == "fp64":
            np.testing.assert_allclose(got[i], want, **TOL64)
        else:
            ok, err = fp32_ok(got[i, 1:], want[1:])
            assert ok.all(), (i, err.max())
            assert abs(got[i, 0] - want[0]) < 1e-3


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_uv_energy_soln12(prec):
    """Solution system: 12-site solute + 1000 waters, 31 A, PME 32^3."""
    print_case(f"uv_energy soln12 {prec}")
    sy = synth.soln12()
    o = O.Oracle(sy.system)
    with ErmGPU(sy.system, prec) as g:
        for f in range(2):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            want = o.uv_energy(sy.solute_mol)
            got = g.get_uv_energy(sy.solute_mol)[0]
            if prec == "fp64":
                np.testing.assert_allclose(got, want, **TOL64)
            else:
                ok, err = fp32_ok(got[1:], want[1:])
                assert ok.all(), err.max()


# ------------------------------------------------------- histogram parity
def test_engconst_refs_exact_counts(tiny):
    """FP64: insdst / edens / ecorr bit-exact against the oracle."""
    print_case("engconst tiny exact counts")
    sy, _ = tiny
    s = sy.system
    o = O.Oracle(s)
    with ErmGPU(s, "fp64", max_batch=7) as g:
        for f in range(3):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            P = poses(sy, 20, 10 + f)
            u = g.engconst(solute_pos=P, want_uv=True)
            for i, sp in enumerate(P):
                want = o.uv_energy(sy.solute_mol, sp)
                np.testing.assert_allclose(u[i], want, **TOL64)
                o.update_histogram(sy.solute_mol, want)
        a, b = o.acc(), g.fetch()
    assert np.array_equal(a["edens"], b["edens"])
    assert np.array_equal(a["ecorr"], b["ecorr"])
    assert a["engnorm"] == b["engnorm"] and a["engsmpl"] == b["engsmpl"] and a["avslf"] == b["avslf"]
    np.testing.assert_allclose(a["minuv"], b["minuv"], rtol=1e-11)
    np.testing.assert_allclose(a["maxuv"], b["maxuv"], rtol=1e-11)


def test_insert_random_matches_oracle(refs):
    """Device-generated poses are the oracle's poses; counts bit-exact."""
    print_case("insert_random refs20")
    sy, pos = refs
    s = sy.system
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    n, seed = 48, 2026
    with ErmGPU(s, "fp64", max_batch=32) as g:
        g.set_frame(pos, sy.cell)
        u = g.insert_random(n, seed, first_index=100, want_uv=True)
        b = g.fetch()
    for i in range(n):
        r, q = O.random_pose(seed, 100 + i)
        sp = O.place_solute(s.bfcoord, s.solute_mass(sy.solute_mol), sy.cell, r, q)
        want = o.uv_energy(sy.solute_mol, sp)
        np.testing.assert_allclose(u[i], want, **TOL64)
        o.update_histogram(sy.solute_mol, want)
    a = o.acc()
    assert np.array_equal(a["edens"], b["edens"])
    assert np.array_equal(a["ecorr"], b["ecorr"])


def test_fp32_error_and_bin_flips(refs):
    """FP32 vs FP64 on 1000 insertions (a frame's batch): every
    per-molecule error within the FP32 bar, and the bin-flip fraction."""
    print_case("fp32 vs fp64 refs20, 1000 insertions")
    sy, pos = refs
    out = {}
    for prec in ("fp64", "fp32"):
        with ErmGPU(sy.system, prec) as g:
            g.set_frame(pos, sy.cell)
            out[prec] = g.insert_random(1000, 7, want_uv=True)
    edges = O.Oracle(sy.system).uvcrd
    _, flips, nent = fp32_report("refs20 1000 insertions", out["fp32"][:, 1:], out["fp64"][:, 1:], edges)
    assert flips / nent < 1e-4


def test_soln_histogram_and_slnuv():
    print_case("engconst soln12")
    sy = synth.soln12()
    s = sy.system
    o = O.Oracle(s)
    with ErmGPU(s, "fp64") as g:
        for f in range(3):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            g.engconst(solutes=[sy.solute_mol])
            o.update_histogram(sy.solute_mol, o.uv_energy(sy.solute_mol))
        a, b = o.acc(), g.fetch()
    assert np.array_equal(a["edens"], b["edens"])
    np.testing.assert_allclose(b["slnuv"][1:], a["slnuv"][1:], rtol=1e-10)
    g2 = None
    del g2


def test_update_histogram_host_energies(tiny):
    """update_histogram alone on energies from the oracle: exact counts."""
    print_case("update_histogram tiny")
    sy, pos = tiny
    s = sy.system
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    P = poses(sy, 10, 21)
    U = np.array([o.uv_energy(sy.solute_mol, sp) for sp in P])
    for u in U:
        o.update_histogram(sy.solute_mol, u)
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        g.update_histogram(U)
        b = g.fetch()
    a = o.acc()
    assert np.array_equal(a["edens"], b["edens"]) and np.array_equal(a["ecorr"], b["ecorr"])


def test_engstore_normalisation(tiny):
    print_case("engstore tiny")
    sy, pos = tiny
    s = sy.system
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        g.insert_random(30, 5)
        raw = g.fetch()
        g.engstore()
        nrm = g.fetch()
    assert np.array_equal(nrm["edens"], raw["edens"] / raw["engnorm"])
    assert np.array_equal(nrm["ecorr"], raw["ecorr"] / raw["engnorm"])


def test_error_codes(tiny):
    """eng_min from an energy below the grid is returned, not aborted."""
    print_case("error codes tiny")
    sy, pos = tiny
    s = sy.system
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        u = np.zeros((1, g.slvmax + 1))
        u[0, 5] = -1e4
        with pytest.raises(ErmError) as ei:
            g.update_histogram(u)
        assert ei.value.code == 1
        cell = np.diag([14.0, 14.0, 14.0])
        cell[0, 2] = 1.0
        with pytest.raises(ErmError) as ei:
            g.set_frame(pos, cell)
        assert ei.value.code == 6


def test_soln300_fp32_large_system():
    """Large solution system: 300-site flexible chain + 20000 waters, 84.5 A,
    PME 96^3 order 6, FP32 pair/PME with FP64 sums, against the FP64 oracle."""
    print_case("soln300 fp32 + fp64")
    sy = synth.soln300()
    s = sy.system
    pos = sy.frame(0)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    want = o.uv_energy(sy.solute_mol)
    res = {}
    for prec in ("fp32", "fp64"):
        with ErmGPU(s, prec) as g:
            g.set_frame(pos, sy.cell)
            res[prec] = g.get_uv_energy(sy.solute_mol)[0]
    np.testing.assert_allclose(res["fp64"], want, **TOL64)
    fp32_report("soln300 vs oracle", res["fp32"][1:], want[1:], o.uvcrd)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_register_path_tunings(refs, prec):
    """Every shape the autotuner may pick for the register path -- x-slab
    and z-column workgroup sizes, the inverse y transform inside the x slab,
    slab depth and planes per group -- gives the oracle's reciprocal
    energies."""
    print_case(f"register-path tunings {prec}")
    sy, pos = refs
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    P = poses(sy, 2, 46)
    want = [o.recpcal(sy.solute_mol, sp) for sp in P]
    seen = set()
    with ErmGPU(sy.system, prec, pme_path="register") as g:
        g.set_frame(pos, sy.cell)
        for st in (256, 512):
            for ct in (256, 512):
                for fy in (False, True):
                    for sp_ in (3, 5, 6, 8, 16):
                        for gp in (4, 8, 9, 16):
                            g.set_tuning(max_batch=0, pme_path="register", slab_planes=sp_, group_planes=gp,
                                         slab_threads=st, column_threads=ct, fold_y=fy)
                            t = g.tuning()
                            seen.add((t["slab_threads"], t["column_threads"], t["fold_y"], t["slab_planes"],
                                      t["group_planes"]))
                            for sp, (we, ws) in zip(P, want):
                                got, gs = g.recpcal(sy.solute_mol, sp)
                                if prec == "fp64":
                                    np.testing.assert_allclose(got, we, **TOL64)
                                    assert gs == pytest.approx(ws, rel=1e-11, abs=1e-9)
                                else:
                                    ok, err = fp32_ok(got, we)
                                    assert ok.all(), (t, err.max())
    assert any(k[2] for k in seen), "the folded y transform was not exercised"
    assert len({k[3] for k in seen}) >= 2


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("which", ["tiny", "refs20"])
def test_pair_paths_agree(tiny, refs, which, prec):
    """Both real-space kernels -- compacted tiles and one thread per
    molecule -- give the oracle's real-space energies (tiny: a box small
    enough that pairs need their own minimum image)."""
    print_case(f"pair paths {which} {prec}")
    sy, pos = tiny if which == "tiny" else refs
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    P = poses(sy, 4, 47)
    want = [o.realcal(sy.solute_mol, sp) for sp in P]
    with ErmGPU(sy.system, prec) as g:
        g.set_frame(pos, sy.cell)
        for path in ("tiles", "molecule"):
            g.set_pair_path(path)
            assert g.pair_path() == path
            for sp, w in zip(P, want):
                got = g.realcal_proc(sy.solute_mol, sp)
                if prec == "fp64":
                    np.testing.assert_allclose(got, w, **TOL64)
                else:
                    ok, err = fp32_ok(got, w)
                    assert ok.all(), (path, err.max())


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("ljswitch", [1, 2, 3])
def test_lj_switches_per_molecule(refs, ljswitch, prec):
    """The potential switch (1), the force switch (2) and the force switch
    with the CHARMM shift (3) (realcal.F90:708-776), per molecule on both
    real-space kernels, against the oracle's (which the brute-force sum and
    the closed forms at lwljcut and upljcut pin); molecules in the switching
    shell [lwljcut, upljcut] are present in every pose."""
    print_case(f"lj switch {ljswitch} {prec}")
    sy, pos = refs
    s = dataclasses.replace(sy.system, ljswitch=ljswitch)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    P = poses(sy, 4, 53)
    want = [o.realcal(sy.solute_mol, sp) for sp in P]
    ref0 = O.Oracle(sy.system)
    ref0.set_frame(pos, sy.cell)
    # the switch changes the energies of some molecules in every pose
    for sp, w in zip(P, want):
        assert np.count_nonzero(np.abs(ref0.realcal(sy.solute_mol, sp) - w) > 1e-9) > 0
    with ErmGPU(s, prec) as g:
        g.set_frame(pos, sy.cell)
        for path in ("tiles", "molecule"):
            g.set_pair_path(path)
            for sp, w in zip(P, want):
                got = g.realcal_proc(sy.solute_mol, sp)
                if prec == "fp64":
                    np.testing.assert_allclose(got, w, **TOL64)
                else:
                    ok, err = fp32_ok(got, w)
                    assert ok.all(), (path, err.max())


@pytest.mark.parametrize("path", ["fused", "unfused", "register"])
def test_pme_paths_agree(refs, path):
    """Every reciprocal-space path (LDS slab; plane + gather; register
    columns) matches the oracle in FP64 and FP32."""
    print_case(f"pme path {path}")
    sy, pos = refs
    o = O.Oracle(sy.system)
    o.set_frame(pos, sy.cell)
    P = poses(sy, 3, 44)
    for prec in ("fp64", "fp32"):
        with ErmGPU(sy.system, prec, pme_path=path) as g:
            g.set_frame(pos, sy.cell)
            for sp in P:
                want, _ = o.recpcal(sy.solute_mol, sp)
                got, _ = g.recpcal(sy.solute_mol, sp)
                if prec == "fp64":
                    np.testing.assert_allclose(got, want, **TOL64)
                else:
                    ok, err = fp32_ok(got, want)
                    assert ok.all(), err.max()


def test_fp32_error_breakdown(refs):
    """Real-space and reciprocal FP32 errors separately (printed)."""
    print_case("fp32 error breakdown")
    sy, pos = refs
    P = poses(sy, 8, 45)
    res = {}
    for prec in ("fp64", "fp32"):
        with ErmGPU(sy.system, prec) as g:
            g.set_frame(pos, sy.cell)
            res[prec] = ([g.realcal_proc(sy.solute_mol, sp) for sp in P],
                         [g.recpcal(sy.solute_mol, sp)[0] for sp in P])
    er = max(np.max(np.abs(a - b)[np.abs(b) <= FP32_EMAX]) for a, b in zip(res["fp32"][0], res["fp64"][0]))
    ek = max(np.max(np.abs(a - b)) for a, b in zip(res["fp32"][1], res["fp64"][1]))
    print(f"fp32 max |dE| real {er:.3e}  recip {ek:.3e}")
    assert er <= FP32_ABS and ek <= FP32_ABS
    for a, b in zip(res["fp32"][0], res["fp64"][0]):
        assert fp32_ok(a, b)[0].all()
