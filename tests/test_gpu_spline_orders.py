"""Spline orders other than the default 4 through the register PME path:
p = 6 takes the z columns' spread-and-FFT kernel and the gather with the
order at compile time; p = 5 takes the spectral z sum and the gather with
the order at run time.  Per-molecule energies of device-generated
insertions against the oracle on the same poses: FP64 within
1e-9 kcal/mol + 1e-11 relative, FP32 within 1e-4 kcal/mol (1e-5 relative
above 100 kcal/mol)."""
import dataclasses

import numpy as np
import pytest

import oracle as O
from bars import FP32_ABS, fp32_ok, fp32_report
from ermod_amd import synth
from ermod_amd.api import ErmGPU

pytestmark = pytest.mark.gpu


def print_case(name):
    print(f"\n[gpu case] {name}", flush=True)


@pytest.mark.parametrize("p", [5, 6])
@pytest.mark.parametrize("prec", ["fp64", "fp32"])
def test_refs20_spline_order(p, prec):
    print_case(f"refs20 spline order {p} {prec}")
    sy = synth.refs20()
    s = dataclasses.replace(sy.system, splodr=p)
    m = sy.solute_mol
    pos = sy.frame(0)
    o = O.Oracle(s)
    o.set_frame(pos, sy.cell)
    n, seed, first = 8, 21, 50
    with ErmGPU(s, prec, max_batch=8) as g:
        g.set_frame(pos, sy.cell)
        u = g.insert_random(n, seed, first_index=first, want_uv=True)
    for i in range(n):
        r, q = O.random_pose(seed, first + i)
        want = o.uv_energy(m, O.place_solute(s.bfcoord, s.solute_mass(m), sy.cell, r, q))
        if prec == "fp64":
            np.testing.assert_allclose(u[i], want, rtol=1e-11, atol=1e-9)
        else:
            ok, err = fp32_ok(u[i, 1:], want[1:])
            assert ok.all(), (i, err.max())
