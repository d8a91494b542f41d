"""Heterogeneous insertion: a mixed solvent (water, 10-site cosolvent, Na+,
Cl-: four species with their own bins) and four rigid test particles of 5,
20, 60 and 150 sites, K = 64.  Per solute size: device random insertions
against the oracle on the same poses (FP64 energies to the usual bar, every
count bit-exact over the four species), FP32 error and bin flips.
"""
import numpy as np
import pytest

import oracle as O
from bars import FP32_ABS, fp32_ok, fp32_report
from ermod_amd import synth
from ermod_amd.api import ErmGPU

pytestmark = pytest.mark.gpu

TOL64 = dict(rtol=1e-11, atol=1e-9)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


@pytest.mark.parametrize("nsite", synth.MIXED_SOLUTE_SITES)
def test_mixed_insertions_fp64(nsite):
    print_case(f"mixed solvent, {nsite}-site solute, fp64")
    sy = synth.mixed(nsite)
    s = sy.system
    pos = sy.frame(0)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    n, seed, first = 12, 404, 5000
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        u = g.insert_random(n, seed, first_index=first, want_uv=True)
        b = g.fetch()
    for i in range(n):
        r, q = O.random_pose(seed, first + i)
        sp = O.place_solute(s.bfcoord, s.solute_mass(sy.solute_mol), sy.cell, r, q)
        want = o.uv_energy(sy.solute_mol, sp)
        np.testing.assert_allclose(u[i], want, **TOL64)
        o.update_histogram(sy.solute_mol, want)
    a = o.acc()
    assert np.array_equal(a["edens"], b["edens"])
    assert np.array_equal(a["ecorr"], b["ecorr"])
    np.testing.assert_allclose(b["minuv"][1:], a["minuv"][1:], rtol=1e-11, atol=1e-9)
    np.testing.assert_allclose(b["maxuv"][1:], a["maxuv"][1:], rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("nsite", synth.MIXED_SOLUTE_SITES)
def test_mixed_fp32_error(nsite):
    """FP32 against FP64 on a frame's share of insertions (1250) for every
    solute size: each of the 8000 molecules' energies within the FP32 bar,
    and the bin flips over the four species' edges."""
    print_case(f"mixed solvent, {nsite}-site solute, fp32 vs fp64, 1250 insertions")
    sy = synth.mixed(nsite)
    pos = sy.frame(1)
    out = {}
    for prec in ("fp64", "fp32"):
        with ErmGPU(sy.system, prec) as g:
            g.set_frame(pos, sy.cell)
            out[prec] = g.insert_random(1250, 9, want_uv=True)
    o = O.Oracle(sy.system)
    spec = sy.system.uvspec[sy.system.physical()]
    offs = np.concatenate([[0], np.cumsum(o.uvmax)])
    e32, e64 = out["fp32"][:, 1:], out["fp64"][:, 1:]
    flips = 0
    for t in range(1, sy.system.numslv + 1):
        edges = o.uvcrd[offs[t - 1]:offs[t]]
        sel = spec == t
        flips += int(np.count_nonzero(np.searchsorted(edges, e32[:, sel], side="left")
                                      != np.searchsorted(edges, e64[:, sel], side="left")))
    fp32_report(f"mixed {nsite}-site, 1250 insertions", e32, e64)
    print(f"[fp32] mixed {nsite}-site bin flips {flips}/{e32.size}", flush=True)
    assert flips / e32.size < 1e-4


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_multi_context_concurrent_matches_sequential(prec):
    """The four solute sizes on one frame: concurrent insertion on four
    streams (erm_insert_random_multi) leaves exactly the accumulators of the
    same insertions run context by context."""
    print_case(f"four solutes concurrently, {prec}")
    from ermod_amd.api import insert_random_multi
    sys_ = [synth.mixed(ns) for ns in synth.MIXED_SOLUTE_SITES]
    pos = sys_[0].frame(2)
    n, seeds, firsts = [40, 40, 24, 24], [3, 4, 5, 6], [0, 100, 200, 300]
    conc = [ErmGPU(sy.system, prec, max_batch=16) for sy in sys_]
    seq = [ErmGPU(sy.system, prec, max_batch=16) for sy in sys_]
    try:
        for sy, a, b in zip(sys_, conc, seq):
            f = sy.frame(2)
            assert np.array_equal(f[: sy.system.natom - sy.system.nsite(sy.solute_mol)],
                                  pos[: sy.system.natom - sy.system.nsite(sy.solute_mol)])
            a.set_frame(f, sy.cell)
            b.set_frame(f, sy.cell)
        insert_random_multi(conc, n, seeds, firsts)
        for k, b in enumerate(seq):
            b.insert_random(n[k], seeds[k], first_index=firsts[k])
        for a, b in zip(conc, seq):
            x, y = a.fetch(), b.fetch()
            for key in ("edens", "ecorr", "minuv", "maxuv"):
                np.testing.assert_array_equal(x[key], y[key])
            assert x["engsmpl"] == y["engsmpl"]
    finally:
        for e in conc + seq:
            e.close()
