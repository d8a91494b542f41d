"""The intermediate-state method on the device (erm_set_intermediate) against
the oracle on the same solution frames: edintm bit-exact, uvintm to 1e-12,
norm_intm / soln_intm / soln_total exact, over two divisions (engclear
keeps soln_intm and soln_total) and through engstore."""
import numpy as np
import pytest

import oracle as O
from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU

pytestmark = pytest.mark.gpu


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_intermediate_matches_oracle(prec):
    print_case(f"intermediate state soln12 {prec}")
    sy = synth.soln12(nwater=400, L=23.0, K=24)
    s = sy.system
    m = sy.solute_mol
    # a bin "empty in the reference solvent": the first one that some but
    # not all of the first division's frames occupy
    o0 = O.Oracle(s)
    hs = []
    for f in range(4):
        o0.set_frame(sy.frame(f), sy.cell)
        hs.append(o0.update_histogram(m, o0.uv_energy(m)) > 0)
    occ = np.sum(hs, axis=0)
    mask = np.zeros(o0.ermax, np.int32)
    mask[np.nonzero((occ >= 1) & (occ <= 3))[0][0]] = 1
    o = O.Oracle(s)
    o.set_intermediate(mask)
    with ErmGPU(s, prec) as g:
        g.set_intermediate(mask)
        for div in range(2):
            o.engclear()
            g.engclear()
            for f in range(4 * div, 4 * div + 4):
                pos = sy.frame(f)
                o.set_frame(pos, sy.cell)
                g.set_frame(pos, sy.cell)
                u = g.engconst(solutes=[m], want_uv=True)[0]
                o.update_histogram(m, u if prec == "fp32" else o.uv_energy(m))
            a = o.intermediate()
            b = g.fetch_intermediate()
            assert np.array_equal(a[0], b[0])
            np.testing.assert_allclose(b[1][1:], a[1][1:], rtol=1e-12)
            assert a[2:] == b[2:], (a[2:], b[2:])
            assert 0 < a[3] < a[4]
        o.engstore()
        g.engstore()
        a, b = o.intermediate(), g.fetch_intermediate()
        np.testing.assert_array_equal(b[0], a[0])


def test_intermediate_needs_a_solution_system():
    print_case("intermediate state refused for test particles")
    sy = synth.tiny_refs()
    with ErmGPU(sy.system, "fp64") as g:
        with pytest.raises(ErmError):
            g.set_intermediate(np.zeros(g.ermax, np.int32))
