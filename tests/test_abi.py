"""The C-ABI library without a GPU: it loads, exports every entry point
include/ermgpu.h declares, its host-only functions agree with the oracle, and
device calls fail with an error code (no CPU fallback)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import oracle as O
from ermod_amd import api as A
from ermod_amd.system import Hist

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "ermgpu.h")).read()
    return sorted(set(re.findall(r"\b(erm_[a-z_0-9]+)\s*\(", src)))


def test_exports_every_declared_symbol():
    names = declared()
    assert len(names) >= 20
    L = A.lib()
    for n in names:
        assert hasattr(L, n), n
    assert set(A.EXPORTS) == set(names)


@pytest.mark.parametrize("h", [Hist(), Hist(pecore=200), Hist(pecore=0), Hist(ecdmin=-40, eccore=30),
                               Hist(eclbin=0.4, ecfbin=0.2, ec0bin=0.05, finfac=4, pecore=30)])
def test_edges_bitwise_vs_oracle(h):
    a, na, sa = A.build_edges(h)
    b, nb, sb = O.build_edges(h)
    assert (na, sa) == (nb, sb) and np.array_equal(a.view(np.uint64), b.view(np.uint64))
    c, nc, _ = A.build_edges(h, self_energy=True)
    assert nc == sa and np.array_equal(c, a[:sa])


def test_edges_error_code():
    with pytest.raises(A.ErmError) as ei:
        A.build_edges(Hist(ecfmns=-19.05))
    assert ei.value.code == 4


def test_hash_vs_oracle():
    v = np.random.default_rng(0).normal(size=123)
    assert A.hash_double(v) == O.hash_double(v)


def test_create_without_device_fails_loudly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except ImportError:
        pass
    p = A.ErmParams()
    p.slttype, p.splodr = 2, 4
    p.ms[:] = [16, 16, 16]
    p.elecut, p.lwljcut, p.upljcut, p.screen, p.block_threshold = 6.0, 5.0, 6.0, 0.5, 3.0
    p.cltype, p.boxshp = 2, 1
    h = C.c_void_p()
    assert A.lib().erm_create(C.byref(p), C.byref(h)) == 12


def test_create_checks_electrostatics_before_the_device():
    """cltype / boxshp consistency as setconf.F90 enforces it: Ewald (1)
    is unsupported, a non-periodic box needs bare Coulomb, soln mode and
    NVT, PPPM an even spline order."""
    def params(**kw):
        p = A.ErmParams()
        p.slttype, p.splodr, p.estype = 1, 4, 1
        p.ms[:] = [16, 16, 16]
        p.elecut, p.lwljcut, p.upljcut, p.screen, p.block_threshold = 6.0, 5.0, 6.0, 0.5, 3.0
        p.cltype, p.boxshp = 2, 1
        for k, v in kw.items():
            setattr(p, k, v)
        return p
    h = C.c_void_p()
    assert A.lib().erm_create(C.byref(params(cltype=1)), C.byref(h)) == 11
    assert A.lib().erm_create(C.byref(params(cltype=5)), C.byref(h)) == 10
    assert A.lib().erm_create(C.byref(params(boxshp=0)), C.byref(h)) == 10
    assert A.lib().erm_create(C.byref(params(boxshp=0, cltype=0, slttype=2)), C.byref(h)) == 10
    assert A.lib().erm_create(C.byref(params(boxshp=0, cltype=0, estype=2)), C.byref(h)) == 10
    assert A.lib().erm_create(C.byref(params(cltype=3, splodr=5)), C.byref(h)) == 10
