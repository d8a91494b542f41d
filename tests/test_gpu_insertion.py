"""Device-generated insertions under the reference's insertion modes
(erm_set_insertion) against the oracle's placement of the same generator
draws: per-molecule FP64 energies to 1e-9 + 1e-11 relative, every count
bit-exact; the aggregate centre follows each frame; argument checks as
setconf.F90:399-432 makes them."""
import numpy as np
import pytest

import oracle as O
from bars import FP32_ABS, fp32_ok, fp32_report
from ermod_amd import synth
from ermod_amd.api import ErmError, ErmGPU
from ermod_amd.system import (INSORG_AGGCEN, INSORG_ORIGIN, INSORG_REFSTR, INSPOS_GAUSS, INSPOS_RANDOM,
                              INSPOS_SLAB_GENERIC, INSPOS_SLAB_SYMMETRIC, INSPOS_SPHERE, INSROT_NOCHANGE,
                              INSROT_RANDOM)

pytestmark = pytest.mark.gpu

TOL64 = dict(rtol=1e-11, atol=1e-9)


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


MODES = [
    ("sphere", (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_RANDOM, 1.0, 5.0)),
    ("sphere_fixed_orientation", (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_NOCHANGE, 0.0, 3.0)),
    ("slab", (INSORG_AGGCEN, INSPOS_SLAB_GENERIC, INSROT_RANDOM, -2.0, 4.0)),
    ("slab_symmetric", (INSORG_AGGCEN, INSPOS_SLAB_SYMMETRIC, INSROT_RANDOM, 1.0, 4.0)),
    ("random_fixed_orientation", (INSORG_ORIGIN, INSPOS_RANDOM, INSROT_NOCHANGE, 0.0, 0.0)),
]


@pytest.mark.parametrize("name,ins", MODES, ids=[m[0] for m in MODES])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_insertion_modes_match_oracle(name, ins, prec):
    print_case(f"insertion mode {name} {prec}")
    sy = synth.refs20(nwater=1500, L=35.6, K=32)
    s = sy.system
    m = sy.solute_mol
    host = [1]
    o = O.Oracle(s)
    n, seed = 40, 77
    with ErmGPU(s, prec, max_batch=16) as g:
        g.set_insertion(*ins, hostspec=host)
        for f in range(2):
            pos = sy.frame(f)
            o.set_frame(pos, sy.cell)
            g.set_frame(pos, sy.cell)
            u = g.insert_random(n, seed, first_index=f * n, want_uv=True)
            cen = O.aggregate_center(s, pos, host)
            for i in range(n):
                sp = O.insertion_pose(s.bfcoord, s.solute_mass(m), sy.cell, ins, seed, f * n + i, cen)
                want = o.uv_energy(m, sp)
                if prec == "fp64":
                    np.testing.assert_allclose(u[i], want, **TOL64)
                    o.update_histogram(m, want)
                else:
                    ok, err = fp32_ok(u[i, 1:], want[1:])
                    assert ok.all(), (i, err.max())
        b = g.fetch()
    if prec == "fp64":
        a = o.acc()
        assert np.array_equal(a["edens"], b["edens"])
        assert np.array_equal(a["ecorr"], b["ecorr"])


def test_sphere_graph_replay_matches_eager():
    print_case("sphere insertion, graph replay vs eager")
    sy = synth.tiny_refs()
    s = sy.system
    ins = (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_RANDOM, 0.5, 4.0)
    with ErmGPU(s, "fp64", max_batch=16) as e, ErmGPU(s, "fp64", max_batch=16) as g:
        for c in (e, g):
            c.set_insertion(*ins, hostspec=[1])
        g.set_tuning(max_batch=16, use_graph=True)
        for f in range(3):
            for c in (e, g):
                c.set_frame(sy.frame(f), sy.cell)
                c.insert_random(48, 5, first_index=48 * f)
            a, b = e.fetch(), g.fetch()
            for k in ("edens", "ecorr", "minuv", "maxuv"):
                np.testing.assert_array_equal(a[k], b[k])


def test_sphere_gradient_poses():
    """erm_solute_gradient's random poses follow the insertion mode."""
    print_case("sphere insertion, gradient poses")
    sy = synth.tiny_refs()
    s = sy.system
    m = sy.solute_mol
    ins = (INSORG_AGGCEN, INSPOS_SPHERE, INSROT_RANDOM, 0.0, 2.0)
    pos = sy.frame(0)
    cen = O.aggregate_center(s, pos, [1])
    confs = np.array([O.insertion_pose(s.bfcoord, s.solute_mass(m), sy.cell, ins, 3, 10 + i, cen)
                      for i in range(6)])
    with ErmGPU(s, "fp64") as g:
        g.set_insertion(*ins, hostspec=[1])
        g.set_frame(pos, sy.cell)
        a = g.solute_gradient(n=6, seed=3, first_index=10)
        b = g.solute_gradient(solute_pos=confs)
    assert a[0] == b[0]
    assert a[1] == pytest.approx(b[1], rel=1e-12, abs=1e-12)
    for x, y in zip(a[2:], b[2:]):
        np.testing.assert_allclose(x, y, rtol=1e-12, atol=1e-12)


def test_insertion_argument_checks():
    print_case("insertion argument checks")
    sy = synth.tiny_refs()
    with ErmGPU(sy.system, "fp64") as g:
        bad = [((INSORG_ORIGIN, INSPOS_SPHERE, 0, 1.0, 2.0, [1]), 10),      # sphere needs AGGCEN
               ((INSORG_AGGCEN, INSPOS_RANDOM, 0, 0.0, 0.0, [1]), 10),      # random needs ORIGIN
               ((INSORG_AGGCEN, INSPOS_SPHERE, 0, 3.0, 2.0, [1]), 10),      # lwreg > upreg
               ((INSORG_AGGCEN, INSPOS_SLAB_SYMMETRIC, 0, -1.0, 2.0, [1]), 10),
               ((INSORG_AGGCEN, INSPOS_SPHERE, 0, 1.0, 2.0, []), 10),       # no host species
               ((INSORG_AGGCEN, INSPOS_SPHERE, 2, 1.0, 2.0, [1]), 10),      # insorient
               ((INSORG_REFSTR, INSPOS_GAUSS, 0, 1.0, 2.0, [1]), 10)]       # no reference structure yet
        for (org, p, rot, lw, up, host), code in bad:
            with pytest.raises(ErmError) as ei:
                g.set_insertion(org, p, rot, lw, up, hostspec=host)
            assert ei.value.code == code, (org, p, rot, lw, up, host)
