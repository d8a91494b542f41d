"""erm_update_histogram_static (the entry point that allocates no device
memory) against erm_update_histogram: the same energies into two contexts
of the same system, two calls in a row (the second smaller, so the static
scratch holds the first call's contents when the second starts), every
accumulator bit for bit; and a call beyond the scratch's capacity is
refused with ERM_E_ARG."""
import numpy as np
import pytest

from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU

pytestmark = pytest.mark.gpu


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


def energies(g, sy, n, seed):
    """Energies of n random insertions (refs) or of the solute n times on
    the frame (soln), from the device path."""
    if sy.system.slttype == 1:
        u = g.get_uv_energy(sy.solute_mol)
        rng = np.random.default_rng(seed)
        # distinct configurations: jitter the molecule energies inside their bins' range
        out = np.repeat(u, n, axis=0)
        out[:, 1:] += rng.uniform(-0.05, 0.05, out[:, 1:].shape)
        return out
    return g.insert_random(n, seed, want_uv=True)


def same(a, b):
    for k in ("edens", "ecorr", "eself", "minuv", "maxuv", "slnuv"):
        if k in a:
            assert np.array_equal(a[k], b[k]), k
    for k in ("engnorm", "engsmpl", "avslf", "voffset"):
        assert a[k] == b[k], k


@pytest.mark.parametrize("kind", ["refs", "soln"])
def test_static_matches_update_histogram(kind):
    print_case(f"update_histogram_static {kind}")
    sy = synth.tiny_refs() if kind == "refs" else synth.tiny_soln()
    s = sy.system
    pos = sy.frame(0)
    with ErmGPU(s, "fp64") as src, ErmGPU(s, "fp64") as a, ErmGPU(s, "fp64") as b:
        for g in (src, a, b):
            g.set_frame(pos, sy.cell, 0.8)
        for n, seed in ((40, 5), (13, 6)):
            u = energies(src, sy, n, seed)
            sol = np.full(n, sy.solute_mol, dtype=np.int32) if kind == "soln" else None
            w = np.random.default_rng(seed).uniform(0.5, 1.5, n)
            a.update_histogram(u, solutes=sol, weights=w)
            b.update_histogram_static(u, solutes=sol, weights=w)
            same(a.fetch(), b.fetch())
        # unit weights
        u = energies(src, sy, 7, 8)
        sol = np.full(7, sy.solute_mol, dtype=np.int32) if kind == "soln" else None
        a.update_histogram(u, solutes=sol)
        b.update_histogram_static(u, solutes=sol)
        same(a.fetch(), b.fetch())


def test_static_refuses_beyond_capacity():
    print_case("update_histogram_static capacity")
    sy = synth.tiny_refs()
    s = sy.system
    with ErmGPU(s, "fp64") as g:
        g.set_frame(sy.frame(0), sy.cell)
        S, E = g.slvmax, g.ermax
        per = 16 * (S + 1) + 8 * (E + 35) + 8
        n = (64 << 20) // per + 1
        u = np.zeros((n, S + 1))
        with pytest.raises(ErmError) as ei:
            g.update_histogram_static(u)
        assert ei.value.code == 10
        assert "exceed" in str(ei.value)
        assert g.fetch()["engsmpl"] == 0.0
