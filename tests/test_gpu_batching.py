"""Graph replay and the autotuner.

* A context in graph mode (random insertions replayed from a captured HIP
  graph) must leave exactly the accumulators an eager context leaves after
  every call of a sequence, with eager calls of its own (uvengy requested,
  which bypasses the graph) and synchronisations in between, across a
  change of frame.
* After autotuning, insertions still match the oracle bit for bit in the
  counts.
"""
import numpy as np
import pytest

import oracle as O
from ermod_amd import synth
from ermod_amd.api import ErmGPU

pytestmark = pytest.mark.gpu


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_graph_replays_match_eager(prec):
    print_case(f"graph replay vs eager {prec}")
    sy = synth.tiny_refs()
    s = sy.system
    frames = [sy.frame(0), sy.frame(1)]
    with ErmGPU(s, prec, max_batch=16) as eager, ErmGPU(s, prec, max_batch=16) as graph:
        graph.set_tuning(max_batch=16, use_graph=True)
        assert graph.tuning()["use_graph"]
        calls = [(0, 40, 7, 0), (0, 40, 7, 40), (0, 40, 8, 0), (1, 40, 7, 80), (1, 40, 9, 120), (1, 40, 9, 160)]
        last_frame = None
        for k, (f, n, seed, first) in enumerate(calls):
            if f != last_frame:
                eager.set_frame(frames[f], sy.cell)
                graph.set_frame(frames[f], sy.cell)
                last_frame = f
            eager.insert_random(n, seed, first_index=first)
            graph.insert_random(n, seed, first_index=first)
            if k == 2:
                # eager work on the graph context between replays
                u1 = graph.insert_random(8, 99, first_index=0, want_uv=True)
                u2 = eager.insert_random(8, 99, first_index=0, want_uv=True)
                np.testing.assert_array_equal(u1, u2)
            a, b = eager.fetch(), graph.fetch()
            for key in ("edens", "ecorr", "minuv", "maxuv"):
                np.testing.assert_array_equal(a[key], b[key], err_msg=f"call {k} {key}")
            for key in ("engnorm", "engsmpl", "avslf"):
                assert a[key] == b[key], (k, key)


def test_autotune_keeps_results(capsys):
    print_case("autotune refs20")
    sy = synth.refs20()
    s = sy.system
    pos = sy.frame(0)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    with ErmGPU(s, "fp64") as g:
        g.set_frame(pos, sy.cell)
        tun, ms = g.autotune(n_trial=512, seed=1)
        print(f"autotune: {tun}, {ms:.3f} ms per 512 insertions", flush=True)
        assert tun["max_batch"] in (64, 128, 256, 512, 1024)
        # accumulators untouched by tuning
        b0 = g.fetch()
        assert b0["engsmpl"] == 0.0 and not b0["edens"].any()
        n, seed = 40, 31
        g.insert_random(n, seed, first_index=7)
        g.insert_random(n, seed, first_index=7 + n)
        b = g.fetch()
    for i in range(2 * n):
        r, q = O.random_pose(seed, 7 + i)
        sp = O.place_solute(s.bfcoord, s.solute_mass(sy.solute_mol), sy.cell, r, q)
        o.update_histogram(sy.solute_mol, o.uv_energy(sy.solute_mol, sp))
    a = o.acc()
    assert np.array_equal(a["edens"], b["edens"])
    assert np.array_equal(a["ecorr"], b["ecorr"])
