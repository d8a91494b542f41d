"""Several test particles in one launch sequence (erm_group_*): the mixed
solvent with its four rigid solutes of 5, 20, 60 and 150 sites.

* Against the oracle (FP64): the same poses through the oracle's
  uv_energy + update_histogram give every member's edens and ecorr bit for
  bit and its extrema to the FP64 bar.
* Against the contexts run alone (FP64 and FP32): every accumulator bit for
  bit, over several batches with uneven shares (one member with none), with
  the z columns and the pairs each as one persistent grid over every solute
  or as one launch per solute.
"""
import numpy as np
import pytest

import oracle as O
from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU, ErmGroup

pytestmark = pytest.mark.gpu

SIZES = synth.MIXED_SOLUTE_SITES


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


@pytest.fixture(scope="module")
def mixed():
    syss = [synth.mixed(ns) for ns in SIZES]
    big = syss[-1]
    return syss, big.frame(3)


def same_acc(a, b):
    for k in ("edens", "ecorr", "minuv", "maxuv"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("engnorm", "engsmpl", "avslf"):
        assert a[k] == b[k], k


def test_group_matches_oracle_fp64(mixed):
    print_case("group of four solutes vs oracle, fp64")
    syss, pos = mixed
    n, seeds, firsts = [6, 5, 4, 3], [21, 22, 23, 24], [0, 10, 20, 30]
    engines = [ErmGPU(sy.system, "fp64") for sy in syss]
    try:
        with ErmGroup(engines) as g:
            g.set_tuning(max_batch=8)          # several batches
            g.set_frame(pos, syss[0].cell)
            g.insert_random(n, seeds, firsts)
            got = [e.fetch() for e in engines]
    finally:
        for e in engines:
            e.close()
    for k, sy in enumerate(syss):
        s = sy.system
        o = O.Oracle(s)
        o.set_frame(pos[:s.natom], sy.cell)
        for i in range(n[k]):
            r, q = O.random_pose(seeds[k], firsts[k] + i)
            sp = O.place_solute(s.bfcoord, s.solute_mass(sy.solute_mol), sy.cell, r, q)
            o.update_histogram(sy.solute_mol, o.uv_energy(sy.solute_mol, sp))
        a = o.acc()
        assert np.array_equal(a["edens"], got[k]["edens"]), SIZES[k]
        assert np.array_equal(a["ecorr"], got[k]["ecorr"]), SIZES[k]
        assert a["engsmpl"] == got[k]["engsmpl"]
        np.testing.assert_allclose(got[k]["minuv"][1:], a["minuv"][1:], rtol=1e-11, atol=1e-9)
        np.testing.assert_allclose(got[k]["maxuv"][1:], a["maxuv"][1:], rtol=1e-11, atol=1e-9)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
@pytest.mark.parametrize("persistent", [0, 1, 2, 3])
def test_group_matches_contexts_alone(mixed, prec, persistent):
    print_case(f"group vs contexts alone, {prec}, persistent {persistent}")
    syss, pos = mixed
    n, seeds, firsts = [300, 0, 170, 233], [5, 6, 7, 8], [0, 0, 1000, 77]
    alone = [ErmGPU(sy.system, prec, max_batch=128) for sy in syss]
    grouped = [ErmGPU(sy.system, prec) for sy in syss]
    try:
        for e, sy, k in zip(alone, syss, range(4)):
            e.set_frame(pos[:sy.system.natom], sy.cell)
            e.insert_random(n[k], seeds[k], first_index=firsts[k])
        with ErmGroup(grouped) as g:
            g.set_tuning(max_batch=256, persistent=persistent)
            g.set_frame(pos, syss[0].cell)
            g.insert_random(n, seeds, firsts)
            # a second call on the same frame accumulates on top
            g.insert_random([7, 9, 0, 5], seeds, [f + 5000 for f in firsts])
            st = g.stage_times()
            assert st["total"] > 0
        for e, sy, k in zip(alone, syss, range(4)):
            e.insert_random([7, 9, 0, 5][k], seeds[k], first_index=firsts[k] + 5000)
        for a, b in zip(alone, grouped):
            same_acc(a.fetch(), b.fetch())
    finally:
        for e in alone + grouped:
            e.close()


def test_group_refuses_mismatched_members(mixed):
    print_case("group refuses different solvents")
    syss, _ = mixed
    other = synth.refs20()
    a = ErmGPU(syss[0].system, "fp64")
    b = ErmGPU(other.system, "fp64")
    try:
        with pytest.raises(ErmError):
            ErmGroup([a, b])
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_group_graph_replay(mixed, prec):
    """Graph replay of the group's batch sequence: three calls of one shape
    over two frames (the first eager and captured, the others replayed),
    with eager calls of the contexts and synchronisations in between; after
    every call each member's accumulators equal the contexts' run alone,
    bit for bit."""
    print_case(f"group graph replay, {prec}")
    syss, _ = mixed
    big = syss[-1]
    frames = [big.frame(3), big.frame(4)]
    n, seeds = [40, 25, 30, 12], [31, 32, 33, 34]
    alone = [ErmGPU(sy.system, prec, max_batch=64) for sy in syss]
    grouped = [ErmGPU(sy.system, prec) for sy in syss]
    other = ErmGPU(syss[1].system, prec)
    try:
        with ErmGroup(grouped) as g:
            g.set_tuning(max_batch=64, persistent=2)
            g.set_graph(True)
            for call, f in enumerate([0, 0, 1, 1]):
                pos = frames[f]
                firsts = [1000 * call + k for k in range(4)]
                g.set_frame(pos, syss[0].cell)
                g.insert_random(n, seeds, firsts)
                # unrelated eager work on the device between replays
                other.set_frame(pos[:syss[1].system.natom], syss[1].cell)
                other.insert_random(17, 5, first_index=call)
                for e, sy, k in zip(alone, syss, range(4)):
                    e.set_frame(pos[:sy.system.natom], sy.cell)
                    e.insert_random(n[k], seeds[k], first_index=firsts[k])
                for a, b in zip(alone, grouped):
                    same_acc(a.fetch(), b.fetch())
                st = g.stage_times()
                assert st["total"] > 0
    finally:
        for e in alone + grouped + [other]:
            e.close()
