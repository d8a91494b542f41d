"""Frames streamed from a DCD file (erm_traj_*) with a PermIndex permutation,
SysWght weights (wrapping around at the end of the file) and skpcnf: the
accumulators equal, bit for bit, those of the same frames handed to
erm_set_frame; the reader reports the end of the file."""
import numpy as np
import pytest

from ermod_amd import synth
from ermod_amd.api import ErmGPU, Trajectory

pytestmark = pytest.mark.gpu


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_trajectory_matches_set_frame(tmp_path, prec):
    print_case(f"DCD trajectory vs set_frame {prec}")
    sy = synth.tiny_refs()
    s = sy.system
    ntraj = s.natom - s.nsite(sy.solute_mol)
    rng = np.random.default_rng(4)
    perm = rng.permutation(ntraj).astype(np.int32)       # trajectory atom i -> atom perm[i]
    nread = 6
    pos = [sy.frame(f).astype(np.float32).astype(np.float64) for f in range(nread)]
    cells = [sy.cell * (1.0 + 0.002 * f) for f in range(nread)]
    traj = np.stack([p[:ntraj][perm] for p in pos])
    path = tmp_path / "sys.dcd"
    synth.write_dcd(path, traj, cells)
    wts = [0.5, 1.25, 2.0, 0.75]
    wpath = tmp_path / "SysWght"
    wpath.write_text("".join(f"{i + 1} {w}\n" for i, w in enumerate(wts)))
    used = [1, 3, 5]                                       # skpcnf = 2: the second of each pair
    wused = [wts[k % len(wts)] for k in used]
    n, seed = 24, 11
    with ErmGPU(s, prec, max_batch=16) as a, ErmGPU(s, prec, max_batch=16) as b:
        with Trajectory(a, path, ntraj, perm=perm, weights=wpath, skpcnf=2) as t:
            for j, k in enumerate(used):
                cell, w = t.next()
                np.testing.assert_array_equal(cell, cells[k])
                assert w == wused[j]
                a.insert_random(n, seed, first_index=n * j)
                b.set_frame(pos[k], cells[k], wused[j])
                b.insert_random(n, seed, first_index=n * j)
            assert t.next() is None
        x, y = a.fetch(), b.fetch()
    for key in ("edens", "ecorr", "minuv", "maxuv"):
        np.testing.assert_array_equal(x[key], y[key])
    for key in ("engnorm", "engsmpl", "avslf"):
        assert x[key] == y[key]
    assert x["engnorm"] == pytest.approx(n * sum(wused), rel=1e-12)
